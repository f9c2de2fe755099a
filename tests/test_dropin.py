"""The C++ drop-in classes (hamiltonian_diagonalization_amd/cpp/hd_dropin.hpp)
driven like the reference's model callers, through the command-line driver
hd_diagonalize: Fock basis, upper-triangle CRS and eigenvalues compared with
the oracle (bit-exact for the basis and the CRS)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import models

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "hamiltonian_diagonalization_amd", "cpp")
DRIVER = os.path.join(ROOT, "hamiltonian_diagonalization_amd", "lib", "hd_diagonalize")


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def run(args, check=True):
    if not os.path.exists(DRIVER):
        pytest.fail("hd_diagonalize is not built (run __graft_entry__.build())")
    p = subprocess.run([DRIVER] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    if check and p.returncode != 0:
        raise AssertionError(f"hd_diagonalize failed ({p.returncode}):\n{p.stdout}\n{p.stderr}")
    return p


def last_json(p):
    return json.loads(p.stdout.strip().splitlines()[-1])


def pseudo_arg(vm):
    return ",".join("%.17g" % v for v in vm)


# ------------------------------------------------------------------ CPU checks

def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "hd_dropin.hpp"\n'
                   "using namespace diagonalization;\n"
                   "int main() {\n"
                   "  utilities::MpiWrapper mpi;\n"
                   "  SpinlessFermionHamiltonian<double> a; SpinlessFermionHamiltonian<dcmplx> b;\n"
                   "  std::function<bool(fock_t)> f = AngularMomentumSector(0, 21);\n"
                   "  return f(0) ? 1 : 0; }\n")
    p = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), "-I" + CPP, str(src)],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


def test_sector_functors_match_oracle_predicate(oracle, tmp_path):
    """The descriptor functors evaluate the reference predicates: compared with the
    oracle's basis (which filters by the restated predicates) over all states."""
    src = tmp_path / "t.cpp"
    src.write_text('#include "hd_dropin.hpp"\n#include <cstdio>\n'
                   "using namespace diagonalization;\n"
                   "int main() {\n"
                   "  AngularMomentumSector lz(2, 9); LinearMomentumSector2D k2(1, 2, 3, 4);\n"
                   "  LinearMomentumSector1D k1(2, 4);\n"
                   "  for (fock_t s = 0; s < (1u << 12); ++s) {\n"
                   "    if (__builtin_popcountll(s) == 4 && s < (1u << 10) && lz(s)) std::printf(\"L %llu\\n\", (unsigned long long)s);\n"
                   "    if (__builtin_popcountll(s) == 4 && k2(s)) std::printf(\"K %llu\\n\", (unsigned long long)s);\n"
                   "    if (__builtin_popcountll(s) == 4 && k1(s)) std::printf(\"O %llu\\n\", (unsigned long long)s);\n"
                   "  }\n  return 0; }\n")
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), "-I" + CPP, str(src),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {t: np.array([int(l.split()[1]) for l in out if l.startswith(t + " ")], dtype=np.uint64)
           for t in "LKO"}
    assert np.array_equal(got["L"], oracle.basis(oracle.sector(oracle.SECTOR_LZ, 2, 9), 4, 10))
    assert np.array_equal(got["K"], oracle.basis(oracle.sector(oracle.SECTOR_K2D, 1, 2, 3, 4), 4, 12))
    assert np.array_equal(got["O"], oracle.basis(oracle.sector(oracle.SECTOR_K1D, 2, 4), 4, 12))


def test_driver_usage_and_empty_sector():
    p = run([], check=False)
    assert p.returncode == 2
    # opposite-parity sector: GenerateFockSpace reports an empty basis (no device work)
    p = run(["--nbr-particles", 3, "--nbr-orbitals", 6, "--lz", 0, "--pseudopotentials", "0,1"])
    assert last_json(p) == {"dim": 0, "empty": True}


# ------------------------------------------------------------------ GPU parity

@pytest.mark.gpu
def test_dropin_sphere_config1(oracle, tmp_path):
    say("drop-in sphere N=8 2S=21 lz=0 Coulomb: basis, upper CRS, 5 eigenvalues")
    vm = models.coulomb_list(21)
    pre = str(tmp_path / "crs")
    p = run(["--nbr-particles", 8, "--nbr-orbitals", 22, "--lz", 0, "--pseudopotentials", pseudo_arg(vm),
             "--nbr-eigenvalues", 5, "--tol", 1e-12, "--dump-crs", pre, "--dump-basis", str(tmp_path / "basis")])
    r = last_json(p)
    say(r)
    bas = np.fromfile(tmp_path / "basis", dtype=np.uint64)
    want_b = oracle.basis(oracle.sector(oracle.SECTOR_LZ, 0, 21), 8, 22)
    assert r["dim"] == 8512 and np.array_equal(bas, want_b)
    kt = hd.sphere_ktable(21)
    vt = hd.sphere_vtable(21, vm)
    up = oracle.build(8, 22, want_b, oracle.LAYOUT_SPHERE, 22, kt, vt, full=False)
    assert np.array_equal(np.fromfile(pre + ".rowstarts", dtype=np.int32), up.rowptr.astype(np.int32))
    assert np.array_equal(np.fromfile(pre + ".cols", dtype=np.int32), up.col)
    assert np.array_equal(np.fromfile(pre + ".values", dtype=np.float64), up.values())
    assert r["nnz_upper"] == len(up.col)
    full = oracle.build(8, 22, want_b, oracle.LAYOUT_SPHERE, 22, kt, vt, full=True)
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    M = sp.csr_matrix((full.values(), full.col, full.rowptr), shape=(8512, 8512))
    want = np.sort(sla.eigsh(M, k=5, which="SA", tol=1e-14, ncv=60)[0])
    got = np.array(r["eigenvalues"])
    assert r["ok"] and r["method"] == "lanczos"
    assert np.max(np.abs(got - want) / np.maximum(1, np.abs(want))) < 1e-10
    assert max(r["residuals"]) <= 1e-12
    # the same solve on the tiled operator (built on first use from the coded values)
    rt = last_json(run(["--nbr-particles", 8, "--nbr-orbitals", 22, "--lz", 0, "--pseudopotentials",
                        pseudo_arg(vm), "--nbr-eigenvalues", 5, "--tol", 1e-12, "--precision", "tiled"]))
    say(rt)
    assert rt["ok"] and rt["method"] == "lanczos" and max(rt["residuals"]) <= 1e-12
    assert np.max(np.abs(np.array(rt["eigenvalues"]) - want) / np.maximum(1, np.abs(want))) < 1e-10


@pytest.mark.gpu
def test_dropin_initial_vector_file(tmp_path):
    """--final-vector writes the ground state in the raw T[nodeDim] format; reading it
    back with --initial-vector (the reference's --use-initial) starts from it."""
    say("drop-in initial/final vector files N=6 2S=15")
    vm = models.coulomb_list(15)
    base = ["--nbr-particles", 6, "--nbr-orbitals", 16, "--lz", 0, "--pseudopotentials", pseudo_arg(vm),
            "--nbr-eigenvalues", 1, "--tol", 1e-11]
    f = str(tmp_path / "gs.bin")
    r0 = last_json(run(base + ["--final-vector", f]))
    assert os.path.getsize(f) == 8 * r0["dim"]
    r1 = last_json(run(base + ["--initial-vector", f]))
    say(r0["matvecs"], "->", r1["matvecs"])
    assert abs(r1["eigenvalues"][0] - r0["eigenvalues"][0]) < 1e-10 * max(1, abs(r0["eigenvalues"][0]))
    assert r1["matvecs"] < r0["matvecs"]
    # a missing file falls back to the random start, as in the reference
    r2 = last_json(run(base + ["--initial-vector", str(tmp_path / "missing.bin")]))
    assert r2["matvecs"] == r0["matvecs"]


@pytest.mark.gpu
def test_dropin_dense_path_small_sector(oracle):
    """dim < 60 goes through FullDiagonalize, as BaseDiagonalize does."""
    say("drop-in dense path N=3 2S=8 lz=4")
    vm = models.coulomb_list(8)
    r = last_json(run(["--nbr-particles", 3, "--nbr-orbitals", 9, "--lz", 4, "--pseudopotentials", pseudo_arg(vm),
                       "--nbr-eigenvalues", 4]))
    assert r["method"] == "full" and r["dim"] < 60
    bas = oracle.basis(oracle.sector(oracle.SECTOR_LZ, 4, 8), 3, 9)
    up = oracle.build(3, 9, bas, oracle.LAYOUT_SPHERE, 9, hd.sphere_ktable(8), hd.sphere_vtable(8, vm))
    want = np.linalg.eigvalsh(up.to_dense(len(bas)))[:4]
    assert np.max(np.abs(np.array(r["eigenvalues"]) - want)) < 1e-11


@pytest.mark.gpu
def test_dropin_flux_lattice_complex_tables(oracle, tmp_path):
    say("drop-in flux lattice 3x4 N=5 complex tables from files")
    kt, vt, eps = models.ofl_tables(3, 4, 11)
    kt.tofile(tmp_path / "k.bin")
    vt.astype(np.complex128).tofile(tmp_path / "v.bin")
    eps.astype(np.complex128).tofile(tmp_path / "e.bin")
    pre = str(tmp_path / "crs")
    r = last_json(run(["--nbr-particles", 5, "--kmax", 12, "--complex", 1, "--ktable", tmp_path / "k.bin",
                       "--vtable", tmp_path / "v.bin", "--eps", tmp_path / "e.bin", "--sector", "k2d:0,0,3,4",
                       "--nbr-eigenvalues", 6, "--tol", 1e-11, "--dense-below", 60, "--dump-crs", pre]))
    say(r)
    bas = oracle.basis(oracle.sector(oracle.SECTOR_K2D, 0, 0, 3, 4), 5, 12)
    up = oracle.build(5, 12, bas, oracle.LAYOUT_OFL, 12, kt, vt, eps=eps, is_complex=True, full=False)
    assert np.array_equal(np.fromfile(pre + ".rowstarts", dtype=np.int32), up.rowptr.astype(np.int32))
    assert np.array_equal(np.fromfile(pre + ".cols", dtype=np.int32), up.col)
    assert np.array_equal(np.fromfile(pre + ".values", dtype=np.complex128), up.values())
    want = np.linalg.eigvalsh(up.to_dense(len(bas)))[:6]
    assert r["ok"] and np.max(np.abs(np.array(r["eigenvalues"]) - want) / np.maximum(1, np.abs(want))) < 1e-10
    # the complex tiled operator through the drop-in
    rt = last_json(run(["--nbr-particles", 5, "--kmax", 12, "--complex", 1, "--ktable", tmp_path / "k.bin",
                        "--vtable", tmp_path / "v.bin", "--eps", tmp_path / "e.bin", "--sector", "k2d:0,0,3,4",
                        "--nbr-eigenvalues", 6, "--tol", 1e-11, "--dense-below", 60, "--precision", "tiled"]))
    say(rt)
    assert rt["ok"] and rt["method"] == "lanczos"
    assert np.max(np.abs(np.array(rt["eigenvalues"]) - want) / np.maximum(1, np.abs(want))) < 1e-10


@pytest.mark.gpu
def test_dropin_two_level(oracle, tmp_path):
    """--nbr-levels 2 (SphereTwoLevelPseudopotentialModel through the drop-in
    TwoLevelFermionFockBasis / TwoLevelSpinlessFermionHamiltonian): basis
    pairs, upper CRS bit-exact against the oracle, eigenvalues against dense,
    and the stored term tables (level 0's k table carries FindM1's label of
    m_maxLz2 + 2)."""
    from hamiltonian_diagonalization_amd import fileio
    n1, N = 6, 5
    norb = 2 * n1 + 2
    say("drop-in two-level n1", n1, "N", N, "lz 1")
    v1 = [0.0, 0.7, 0.0, 0.2]
    v2 = [0.0, 0.5, 0.0, 0.3, 0.0, 0.1]
    e = np.linspace(-0.3, 0.4, norb)
    pre = str(tmp_path / "crs")
    out = tmp_path / "out"
    out.mkdir()
    p = run(["--nbr-levels", 2, "--nbr-particles", N, "--nbr-orbitals", norb, "--lz", 1,
             "--pseudopotentials", pseudo_arg(v1), "--pseudopotentials-2ll", pseudo_arg(v2),
             "--energy-levels", pseudo_arg(e), "--nbr-eigenvalues", 3, "--tol", 1e-12, "--dump-crs", pre,
             "--dump-basis", str(tmp_path / "basis"), "--store-terms", 1, "--file-format", 1, "--out-path", str(out)])
    r = last_json(p)
    say(r)
    f, g = oracle.two_level_basis(N, n1, 1)
    assert r["dim"] == len(f)
    assert np.array_equal(np.fromfile(str(tmp_path / "basis") + ".first", dtype=np.uint64), f)
    assert np.array_equal(np.fromfile(str(tmp_path / "basis") + ".second", dtype=np.uint64), g)
    kt1, kt2 = oracle.two_level_ktable(n1, 0), oracle.two_level_ktable(n1, 1)
    vt1, vt2 = hd.sphere_vtable(n1 - 1, v1), hd.sphere_vtable(n1 + 1, v2)
    up = oracle.two_level_build(n1, f, g, kt1, vt1, kt2, vt2, e[:n1], e[n1:], full=False)
    assert np.array_equal(np.fromfile(pre + ".rowstarts", dtype=np.int32), up.rowptr.astype(np.int32))
    assert np.array_equal(np.fromfile(pre + ".cols", dtype=np.int32), up.col)
    assert np.array_equal(np.fromfile(pre + ".values", dtype=np.float64), up.values())
    w = np.linalg.eigvalsh(up.to_dense(len(f), mirror=True))
    got = np.array(r["eigenvalues"])
    assert r["ok"] and abs(got[0] - w[0]) < 1e-9
    assert all(np.min(np.abs(w - x)) < 1e-9 for x in got)
    for tag, kt, vt, L in (("LLL", kt1, vt1, n1), ("2LL", kt2, vt2, n1 + 2)):
        kmax, v, k = fileio.read_term_table(str(out / f"quartic_coefficient_table_{tag}_L_{norb}.dat"), "sphere")
        assert kmax == L and np.array_equal(k, kt) and np.array_equal(v, vt)
    assert 0xFFFE in kt1 and 0xFFFF not in kt1


@pytest.mark.gpu
def test_dropin_flux_lattice_from_bloch(oracle, tmp_path):
    """--bloch: the driver generates the quartic table from Bloch coefficients
    (Generate4KTable + GenerateQuarticTerms + SetCoefficient(0.5 * strength))
    and builds; the upper CRS matches the oracle's on the oracle's V table."""
    dim_x, dim_y, cut, n, strength = 3, 4, 2, 4, 1.7
    K = dim_x * dim_y
    say("drop-in flux lattice from Bloch coefficients 3x4 N=4 cut 2")
    rng = np.random.default_rng(11)
    b = rng.normal(size=(K, (2 * cut + 1) ** 2, 2)) + 1j * rng.normal(size=(K, (2 * cut + 1) ** 2, 2))
    b /= np.linalg.norm(b.reshape(K, -1), axis=1)[:, None, None]
    b.astype(np.complex128).tofile(tmp_path / "bloch.bin")
    eps = np.linspace(-0.5, 0.5, K).astype(np.complex128)
    eps.tofile(tmp_path / "e.bin")
    pre = str(tmp_path / "crs")
    p = run(["--nbr-particles", n, "--kmax", K, "--layout", "ofl", "--complex", 1, "--dim-x", dim_x, "--dim-y", dim_y,
             "--bloch", tmp_path / "bloch.bin", "--cutoff", cut, "--interaction-strength", strength,
             "--eps", tmp_path / "e.bin", "--sector", f"k2d:1,2,{dim_x},{dim_y}", "--nbr-eigenvalues", 2,
             "--dense-below", 10, "--dump-crs", pre])
    r = last_json(p)
    v = oracle.ofl_vtable(dim_x, dim_y, cut, cut, b)
    c = 0.5 * strength  # complex product with (c, 0): (a c - b 0, a 0 + b c)
    vt = (v.real * c - v.imag * 0.0) + 1j * (v.real * 0.0 + v.imag * c)
    kt = hd.ofl_ktable(dim_x, dim_y)
    states = oracle.basis(oracle.sector(oracle.SECTOR_K2D, 1, 2, dim_x, dim_y), n, K)
    up = oracle.build(n, K, states, oracle.LAYOUT_OFL, K, kt, vt, eps=eps, is_complex=True, full=False)
    assert r["dim"] == len(states)
    assert np.array_equal(np.fromfile(pre + ".cols", dtype=np.int32), up.col)
    assert np.array_equal(np.fromfile(pre + ".values", dtype=np.complex128), up.values())


def test_hash_table_file_format(tmp_path):
    """QuarticTermHashTables / QuadraticTermHashTables files: the reference's
    layout (term_hash_tables_base.hpp:139-219, multi_key_hash.hpp:480-669):
    header, then the k1 multimap (uint32 size, uint64 keys Pack4x16(k2,k3,k4,0),
    uint16 k1), then the value map (uint32 size, uint64 keys
    Pack4x16(k1,k2,k3,k4), complex values); a written file reads back to the
    same tables, in binary and in text."""
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include "hd_dropin.hpp"
using namespace diagonalization;
int main(int argc, char** argv) {
  utilities::MpiWrapper mpi;
  const kState_t K = 6;
  OflQuarticTermTables arr; DiagonalQuadraticTermTables<dcmplx> quad;
  arr.Initialize(K, mpi); quad.Initialize(K, mpi);
  for (kState_t k4 = 0; k4 < K; ++k4) for (kState_t k3 = k4; k3 < K; ++k3) for (kState_t k2 = 0; k2 < K; ++k2) {
    const kState_t k1 = (kState_t)((k3 + k4 + 2 * K - k2) % K);
    arr.SetK1(k1, k2, k3, k4);
    arr.SetVkkkk(dcmplx(0.125 * k1 + k2, 0.5 * k3 - k4 + 0.1), k1, k2, k3, k4); }
  for (kState_t k = 0; k < K; ++k) quad.SetEkk(dcmplx(-0.25 * k, 0.0), k);
  QuarticTermHashTables hq; QuadraticTermHashTables hd2;
  hq.Initialize(K); hd2.Initialize(K);
  hq.SetQuarticFromArray(&arr); hd2.SetQuadraticFromArray(&quad);
  const std::string dir = argv[1];
  for (int f = 0; f < 2; ++f) {
    const io::fileFormat_t fmt = f ? io::_BINARY_ : io::_TEXT_;
    const std::string tag = f ? "bin" : "txt";
    hq.ToFile(dir + "/quartic." + tag, fmt, mpi); hd2.ToFile(dir + "/quadratic." + tag, fmt, mpi);
    QuarticTermHashTables rq; QuadraticTermHashTables rd;
    rq.FromFile(dir + "/quartic." + tag, fmt, mpi); rd.FromFile(dir + "/quadratic." + tag, fmt, mpi);
    rq.ToFile(dir + "/quartic2." + tag, fmt, mpi); rd.ToFile(dir + "/quadratic2." + tag, fmt, mpi);
    kState_t a[8], b[8]; iSize_t na = 0, nb = 0;
    for (kState_t k2 = 0; k2 < K; ++k2) for (kState_t k3 = 0; k3 < K; ++k3) for (kState_t k4 = 0; k4 < K; ++k4) {
      hq.GetK1(a, na, k2, k3, k4); rq.GetK1(b, nb, k2, k3, k4);
      if (na != 1 || nb != 1 || a[0] != b[0]) return 1;
      if (std::abs(hq.GetVkkkk(a[0], k2, k3, k4) - rq.GetVkkkk(b[0], k2, k3, k4)) > 1e-13) return 2;
      if (f && hq.GetVkkkk(a[0], k2, k3, k4) != arr.GetVkkkk(a[0], k2, k3, k4)) return 3; }
    for (kState_t k = 0; k < K; ++k) { rd.GetK1(a, na, k); if (na != 1 || a[0] != k || rd.GetEkk(k, k) != quad.GetEkk(k, k)) return 4; }
  }
  return 0; }
''')
    exe = tmp_path / "t"
    p = subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), "-I" + CPP, str(src),
                        "-o", str(exe)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    p = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True)
    assert p.returncode == 0, (p.returncode, p.stdout, p.stderr)
    header = b"# File contains term table data in hash table format"
    K = 6
    for name in ("quartic", "quadratic"):
        assert (tmp_path / f"{name}.bin").read_bytes() == (tmp_path / f"{name}2.bin").read_bytes()
        assert (tmp_path / f"{name}.txt").read_text() == (tmp_path / f"{name}2.txt").read_text()
        raw = (tmp_path / f"{name}.bin").read_bytes()
        n = int(np.frombuffer(raw[:4], dtype=np.uint32)[0])
        assert raw[4:4 + n] == header
        off = 4 + n
        size = int(np.frombuffer(raw[off:off + 4], dtype=np.uint32)[0])
        keys = np.frombuffer(raw[off + 4:off + 4 + 8 * size], dtype=np.uint64)
        k1 = np.frombuffer(raw[off + 4 + 8 * size:off + 4 + 10 * size], dtype=np.uint16)
        off += 4 + 10 * size
        vsize = int(np.frombuffer(raw[off:off + 4], dtype=np.uint32)[0])
        vkeys = np.frombuffer(raw[off + 4:off + 4 + 8 * vsize], dtype=np.uint64)
        vals = np.frombuffer(raw[off + 4 + 8 * vsize:off + 4 + 24 * vsize], dtype=np.complex128)
        assert off + 4 + 24 * vsize == len(raw)
        if name == "quartic":
            assert size == vsize == K ** 3
            k2, k3, k4 = (keys >> 48) & 0xFFFF, (keys >> 32) & 0xFFFF, (keys >> 16) & 0xFFFF
            assert np.all(keys & 0xFFFF == 0)
            assert np.array_equal(k1, (k3 + k4 + 2 * K - k2) % K)
            vk1, vk2, vk3, vk4 = (vkeys >> 48) & 0xFFFF, (vkeys >> 32) & 0xFFFF, (vkeys >> 16) & 0xFFFF, vkeys & 0xFFFF
            assert np.array_equal(vk1, (vk3 + vk4 + 2 * K - vk2) % K)
            # GetVkkkk of the folded array: k4 > k3 reads the (k3, k4) entry with k1 for k2
            want = np.where(vk4 > vk3, 0.125 * vk2 + vk1 + 1j * (0.5 * vk4 - vk3 + 0.1),
                            0.125 * vk1 + vk2 + 1j * (0.5 * vk3 - vk4 + 0.1))
            assert np.array_equal(vals, want)
        else:
            assert size == vsize == K
            assert np.array_equal(keys, np.arange(K)) and np.array_equal(k1, np.arange(K))
            assert np.array_equal(vkeys, (np.arange(K, dtype=np.uint64) << np.uint64(32)) | np.arange(K, dtype=np.uint64))
            assert np.array_equal(vals, -0.25 * np.arange(K) + 0j)


def _ofl_files(tmp_path, kt, vt, eps):
    kt.tofile(tmp_path / "k.bin")
    vt.astype(np.complex128).tofile(tmp_path / "v.bin")
    eps.astype(np.complex128).tofile(tmp_path / "e.bin")
    return ["--ktable", tmp_path / "k.bin", "--vtable", tmp_path / "v.bin", "--eps", tmp_path / "e.bin"]


@pytest.mark.gpu
def test_dropin_use_hash_matches_array_path(oracle, tmp_path):
    """--use-hash 1 (run_test.sh:7-9): the hash tables of ConvertTableFormat
    build, through the general device build, the same upper CRS as the array
    tables, bit for bit; stored and retrieved in the hash-table file format
    (binary) they build it again."""
    X, Y, n = 3, 4, 5
    K = X * Y
    kt, vt, eps = models.ofl_tables(X, Y, 11)
    files = _ofl_files(tmp_path, kt, vt, eps)
    out = tmp_path / "out"
    out.mkdir()
    base = ["--nbr-particles", n, "--kmax", K, "--layout", "ofl", "--complex", 1, "--dim-x", X, "--dim-y", Y,
            "--sector", "k2d:0,0,3,4", "--nbr-eigenvalues", 5, "--tol", 1e-11, "--dense-below", 60,
            "--out-path", str(out), "--file-format", 1]
    say("drop-in --use-hash 1, 3x4 N=5")
    bas = oracle.basis(oracle.sector(oracle.SECTOR_K2D, 0, 0, X, Y), n, K)
    up = oracle.build(n, K, bas, oracle.LAYOUT_OFL, K, kt, vt, eps=eps, is_complex=True, full=False)
    want = np.linalg.eigvalsh(up.to_dense(len(bas)))[:5]
    for extra in (["--store-terms", 1], ["--retrieve-terms", 1]):
        pre = str(tmp_path / ("crs" + extra[0]))
        r = last_json(run(base + files + ["--use-hash", 1, "--dump-crs", pre] + extra))
        say(extra, r)
        assert np.array_equal(np.fromfile(pre + ".rowstarts", dtype=np.int32), up.rowptr.astype(np.int32))
        assert np.array_equal(np.fromfile(pre + ".cols", dtype=np.int32), up.col)
        assert np.array_equal(np.fromfile(pre + ".values", dtype=np.complex128), up.values())
        assert r["ok"] and np.max(np.abs(np.array(r["eigenvalues"]) - want)) < 1e-10
    assert (out / f"quartic_coefficient_table_kx_{X}_ky_{Y}.dat").exists()


@pytest.mark.gpu
def test_dropin_wannier_basis(oracle, tmp_path):
    """--basis 1 (run_test.sh:19): the hybrid Wannier tables built through the
    drop-in give the upper CRS of the oracle's general build on hd_ofl_wannier's
    lists (--wannier-matrix), and with the change of basis generated from
    Bloch vectors (--bloch) the lowest eigenvalues of the k_y sector are the
    lowest of the Bloch sectors (k_x, k_y) together (offset 0, E(k_x) = E(-k_x))."""
    X, Y, n, ky = 3, 4, 4, 1
    K = X * Y
    rng = np.random.default_rng(21)
    E = rng.uniform(-0.3, 0.3, size=(X, Y))
    E = (E + E[(-np.arange(X)) % X]) / 2
    eps = E.reshape(K).astype(np.complex128)
    eps.tofile(tmp_path / "e.bin")
    # (1) raw tables and an explicit change of basis: bit-exact CRS
    kt, vt, _ = models.ofl_tables(X, Y, 4)
    U = models.wannier_matrix(X, Y, 0.3, np.exp(2j * np.pi * rng.random((X, Y))))
    U.astype(np.complex128).tofile(tmp_path / "u.bin")
    kt.tofile(tmp_path / "k.bin")
    vt.astype(np.complex128).tofile(tmp_path / "v.bin")
    pre = str(tmp_path / "crs")
    say("drop-in --basis 1 with --wannier-matrix, 3x4 N=4")
    r = last_json(run(["--nbr-particles", n, "--kmax", K, "--layout", "ofl", "--complex", 1, "--dim-x", X,
                       "--dim-y", Y, "--ktable", tmp_path / "k.bin", "--vtable", tmp_path / "v.bin",
                       "--eps", tmp_path / "e.bin", "--basis", 1, "--kx-shift", 0.3, "--tol-vkkkk", 1e-12,
                       "--wannier-matrix", tmp_path / "u.bin", "--sector", f"k1d:{ky},{Y}", "--nbr-eigenvalues", 4,
                       "--dense-below", 0, "--dump-crs", pre]))
    say(r)
    lists = hd.ofl_wannier(X, Y, 0.3, kt, vt, eps, U, 1e-12)
    bas = oracle.basis(oracle.sector(oracle.SECTOR_K1D, ky, Y), n, K)
    up = oracle.build_general(K, bas, K, full=False, **lists)
    assert np.array_equal(np.fromfile(pre + ".rowstarts", dtype=np.int32), up.rowptr.astype(np.int32))
    assert np.array_equal(np.fromfile(pre + ".cols", dtype=np.int32), up.col)
    assert np.array_equal(np.fromfile(pre + ".values", dtype=np.complex128), up.values())
    # (2) Bloch coefficients: V and the change of basis both from them
    cut = 2
    b = rng.normal(size=(K, (2 * cut + 1) ** 2, 2)) + 1j * rng.normal(size=(K, (2 * cut + 1) ** 2, 2))
    b /= np.linalg.norm(b.reshape(K, -1), axis=1)[:, None, None]
    b.astype(np.complex128).tofile(tmp_path / "bloch.bin")
    common = ["--nbr-particles", n, "--kmax", K, "--layout", "ofl", "--complex", 1, "--dim-x", X, "--dim-y", Y,
              "--bloch", tmp_path / "bloch.bin", "--cutoff", cut, "--eps", tmp_path / "e.bin",
              "--nbr-eigenvalues", 4, "--tol", 1e-11, "--dense-below", 0]
    say("drop-in --basis 1 from Bloch vectors")
    w = last_json(run(common + ["--basis", 1, "--sector", f"k1d:{ky},{Y}"]))
    bloch = []
    for kx in range(X):
        rb = last_json(run(common + ["--sector", f"k2d:{kx},{ky},{X},{Y}"]))
        assert rb["ok"]
        bloch.extend(rb["eigenvalues"])
    say("wannier", w["eigenvalues"], "bloch", sorted(bloch)[:4])
    assert w["ok"] and np.max(np.abs(np.array(w["eigenvalues"]) - np.sort(bloch)[:4])) < 1e-9


@pytest.mark.gpu
def test_dropin_opaque_predicates(oracle, tmp_path):
    """GenerateFockSpace with the std::function the reference's callers pass
    (pseudopotential_model.cpp:185-187): a bound L_z test is recognised as the
    device sector (the same upper CRS as --lz, and a hint); a predicate no
    sector reproduces is enumerated on the host and built by the general
    build, bit for bit the oracle's over that state list."""
    n, two_s = 6, 15
    K = two_s + 1
    vm = [0.0, 1.0, 0.0, 0.45]
    common = ["--nbr-particles", n, "--nbr-orbitals", K, "--pseudopotentials", pseudo_arg(vm),
              "--nbr-eigenvalues", 3, "--tol", 1e-12, "--dense-below", 10]
    say("drop-in opaque predicates, N=6 2S=15")
    p0 = run(common + ["--lz", 0, "--dump-crs", str(tmp_path / "a")])
    p1 = run(common + ["--sector", f"bind-lz:0,{two_s}", "--dump-crs", str(tmp_path / "b")])
    assert "AngularMomentumSector" in p1.stderr
    for ext in (".rowstarts", ".cols", ".values"):
        assert (tmp_path / ("a" + ext)).read_bytes() == (tmp_path / ("b" + ext)).read_bytes()
    # the product's schedule is autotuned per run: equal to rounding
    assert np.max(np.abs(np.array(last_json(p0)["eigenvalues"]) - np.array(last_json(p1)["eigenvalues"]))) < 1e-12
    mask = 0b1001
    p2 = run(common + ["--sector", f"exclude:0,{two_s},{mask}", "--dump-crs", str(tmp_path / "c"),
                       "--dump-basis", str(tmp_path / "basis")])
    assert "general build" in p2.stderr
    r2 = last_json(p2)
    full = oracle.basis(oracle.sector(oracle.SECTOR_LZ, 0, two_s), n, K)
    bas = full[(full & mask) == 0]
    assert np.array_equal(np.fromfile(str(tmp_path / "basis"), dtype=np.uint64), bas)
    up = oracle.build(n, K, bas, oracle.LAYOUT_SPHERE, K, hd.sphere_ktable(two_s), hd.sphere_vtable(two_s, vm))
    assert np.array_equal(np.fromfile(str(tmp_path / "c.rowstarts"), dtype=np.int32), up.rowptr.astype(np.int32))
    assert np.array_equal(np.fromfile(str(tmp_path / "c.cols"), dtype=np.int32), up.col)
    assert np.array_equal(np.fromfile(str(tmp_path / "c.values"), dtype=np.float64), up.values())
    w = np.linalg.eigvalsh(up.to_dense(len(bas)))[:3]
    assert r2["ok"] and np.max(np.abs(np.array(r2["eigenvalues"]) - w)) < 1e-10
