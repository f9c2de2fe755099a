"""The reference's on-disk formats: CRS matrices, term tables, eigensystems and
ARPACK start vectors, in text and binary.

The byte layouts are checked field by field against the write sequences in
crs_sparse_matrix.hpp:662-751, term_tables_base.hpp:179-345 and
hamiltonian_base.hpp:1204-1432.  The numpy implementation
(hamiltonian_diagonalization_amd.fileio) and the C++ drop-in classes
(cpp/hd_dropin.hpp) are written separately from those layouts; files written by
one and read and rewritten by the other must come back byte-identical, text
files included (15 significant digits on both sides).  The GPU cases run the
drop-in driver with the reference's file options and compare its files with
the oracle."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import hamiltonian_diagonalization_amd as hd
from hamiltonian_diagonalization_amd import fileio, models

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "hamiltonian_diagonalization_amd", "cpp")
LIB = os.path.join(ROOT, "hamiltonian_diagonalization_amd", "lib")
DRIVER = os.path.join(LIB, "hd_diagonalize")
HEADER = b"# File contains term table data in array format"


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _random_crs(rng, rows, complex_values=False):
    counts = rng.integers(0, 6, size=rows)
    counts[-1] = 3  # the last row holds entries (the text layout omits its end)
    rowptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(rows, size=c, replace=False)) for c in counts]).astype(np.int32)
    val = rng.normal(size=rowptr[-1]) * 10.0 ** rng.integers(-8, 4, size=rowptr[-1])
    if complex_values:
        val = val + 1j * rng.normal(size=rowptr[-1])
    return rowptr, col, val


# ------------------------------------------------------------------ byte layouts

def test_crs_binary_byte_layout(tmp_path):
    rowptr, col, val = _random_crs(np.random.default_rng(0), 17)
    p = tmp_path / "m.bin"
    fileio.write_crs(p, rowptr, col, val, fileio.BINARY)
    raw = p.read_bytes()
    nnz, rows = len(val), 17
    # uint32 nnz, uint32 rows, int32 rowStarts[rows+1], int32 cols[nnz], double data[nnz]
    assert len(raw) == 8 + 4 * (rows + 1) + 4 * nnz + 8 * nnz
    assert struct.unpack_from("<II", raw, 0) == (nnz, rows)
    assert list(struct.unpack_from("<%di" % (rows + 1), raw, 8)) == rowptr.tolist()
    assert list(struct.unpack_from("<%di" % nnz, raw, 8 + 4 * (rows + 1))) == col.tolist()
    assert list(struct.unpack_from("<%dd" % nnz, raw, 8 + 4 * (rows + 1) + 4 * nnz)) == val.tolist()


def test_crs_text_layout(tmp_path):
    rowptr = np.array([0, 2, 2, 3])
    col = np.array([0, 2, 1], dtype=np.int32)
    val = np.array([1.5, -1.0 / 3.0, 2e-20])
    p = tmp_path / "m.txt"
    fileio.write_crs(p, rowptr, col, val, fileio.TEXT)
    # nnz, rows, rowStarts[0..rows-1], then "i col value" at setprecision(15)
    assert p.read_text() == "3\n3\n0\n2\n2\n0 0 1.5\n0 2 -0.333333333333333\n2 1 2e-20\n"
    zr, zc, zv = fileio.read_crs(p, fileio.TEXT)
    assert np.array_equal(zr, rowptr) and np.array_equal(zc, col)
    assert np.allclose(zv, val, rtol=1e-14, atol=0)


def test_term_table_binary_byte_layout(tmp_path):
    kmax = 6
    kt = hd.sphere_ktable(kmax - 1)
    vt = hd.sphere_vtable(kmax - 1, models.coulomb_list(kmax - 1))
    p = tmp_path / "q.bin"
    fileio.write_term_table(p, "sphere", kmax, vt, kt, fileio.BINARY)
    raw = p.read_bytes()
    n = kmax ** 3
    hl = len(HEADER)
    assert hl == 47
    assert len(raw) == 4 + hl + 4 + 4 + 2 * n + 8 * n
    assert struct.unpack_from("<I", raw, 0)[0] == hl and raw[4:4 + hl] == HEADER
    assert struct.unpack_from("<I", raw, 4 + hl)[0] == 4  # four labels
    # kMax: the 2-byte value in a 4-byte field
    assert struct.unpack_from("<HH", raw, 8 + hl) == (kmax, 0)
    assert np.array_equal(np.frombuffer(raw, dtype="<u2", count=n, offset=12 + hl), kt)
    assert np.array_equal(np.frombuffer(raw, dtype="<f8", count=n, offset=12 + hl + 2 * n), vt)
    # a reader ignores the upper two bytes of the kMax field (padding in the reference)
    bad = bytearray(raw)
    bad[10 + hl:12 + hl] = b"\xab\xcd"
    p2 = tmp_path / "q2.bin"
    p2.write_bytes(bytes(bad))
    k2, v2, t2 = fileio.read_term_table(p2, "sphere", fileio.BINARY)
    assert k2 == kmax and np.array_equal(v2, vt) and np.array_equal(t2, kt)


def test_quadratic_table_layouts(tmp_path):
    e = np.array([0.25, -1.0, 1.0 / 7.0])
    p = tmp_path / "e.bin"
    fileio.write_term_table(p, "quadratic", 3, e, fmt=fileio.BINARY)
    raw = p.read_bytes()
    assert len(raw) == 4 + 47 + 8 + 8 * 3  # two labels: no k table
    assert struct.unpack_from("<II", raw, 51) == (2, 3)
    t = tmp_path / "e.txt"
    fileio.write_term_table(t, "quadratic", 3, e, fmt=fileio.TEXT)
    assert t.read_text() == HEADER.decode() + "\n2\n3\n0\t0\t0.25\n1\t1\t-1\n2\t2\t0.142857142857143\n"


def test_eigensystem_layouts(tmp_path):
    ev = np.array([-1.25, 0.5])
    vec = np.array([[0.6, 0.8, 0.0], [0.0, 0.0, -1.0]])
    p = tmp_path / "s.bin"
    fileio.write_eigensystem(p, ev, vec, fileio.BINARY)
    raw = p.read_bytes()
    assert len(raw) == 4 + 16 + 4 + 48
    assert struct.unpack_from("<I2dI", raw, 0) == (2, -1.25, 0.5, 3)
    t = tmp_path / "s.txt"
    fileio.write_eigensystem(t, ev, vec, fileio.TEXT)
    assert t.read_text() == ("##  NBR EIGENVALUES\n2\n##  EIGENVALUES\n-1.25\n0.5\n"
                             "##  FOCK SPACE DIMENSION\n3\n##  EIGENVECTORS\n0.6\n0.8\n0\n0\n0\n-1\n")
    for f, fmt in ((p, fileio.BINARY), (t, fileio.TEXT)):
        e2, v2 = fileio.read_eigensystem(f, fmt)
        assert np.array_equal(e2, ev) and np.array_equal(v2, vec)
    # values only
    fileio.write_eigensystem(p, ev, None, fileio.BINARY)
    assert p.stat().st_size == 20
    assert fileio.read_eigensystem(p, fileio.BINARY)[1] is None


def test_initial_vector_is_raw(tmp_path):
    v = np.random.default_rng(1).normal(size=11)
    p = tmp_path / "init.bin"
    fileio.write_initial_vector(p, v)
    assert p.read_bytes() == v.astype("<f8").tobytes()
    assert np.array_equal(fileio.read_initial_vector(p), v)


@pytest.mark.parametrize("fmt", [fileio.BINARY, fileio.TEXT])
def test_numpy_round_trips(tmp_path, fmt):
    rng = np.random.default_rng(2)
    tol = 0 if fmt == fileio.BINARY else 1e-14
    for cplx in (False, True):
        rowptr, col, val = _random_crs(rng, 23, cplx)
        p = tmp_path / ("c%d" % cplx)
        fileio.write_crs(p, rowptr, col, val, fmt)
        r, c, v = fileio.read_crs(p, fmt, complex_values=cplx)
        assert np.array_equal(r, rowptr) and np.array_equal(c, col)
        assert np.all(np.abs(v - val) <= tol * np.abs(val))
    kt, vt, eps = models.ofl_tables(3, 4, 5)
    p = tmp_path / "ofl"
    fileio.write_term_table(p, "ofl", 12, vt, kt, fmt)
    k, v, t = fileio.read_term_table(p, "ofl", fmt, complex_values=True)
    assert k == 12
    # the text layout walks all (k2, k3, k4) through GetK1/GetVkkkk; reading it
    # back through SetK1/SetVkkkk restores every slot those accessors reach
    reach = np.zeros(len(vt), dtype=bool)
    for k4 in range(12):
        for k3 in range(12):
            for k2 in range(12):
                r0 = fileio._fold(12, k3, k4) * 12
                reach[r0 + (int(kt[r0 + k2]) if k4 > k3 else k2)] = True
    if fmt == fileio.BINARY:
        reach[:] = True
    assert np.array_equal(t, kt)
    assert np.all(np.abs(v[reach] - vt[reach]) <= tol * np.abs(vt[reach]))


# ------------------------------------------------------------------ C++ interop

CPP_PROGRAM = r'''
#include "hd_dropin.hpp"
using namespace diagonalization;
int main(int argc, char** argv)
{
    utilities::MpiWrapper mpi;
    const std::string d = argv[1];
    const std::string mode = argc > 2 ? argv[2] : "";
    if (mode == "bad-header") {
        SphereQuarticTermTables t;
        t.FromFile(d + "/bad.dat", io::_BINARY_, mpi);
        return 0;
    }
    if (mode == "missing") {
        utilities::CrsSparseMatrix<double> c;
        c.FromFile(d + "/does_not_exist.dat", io::_TEXT_, mpi);
        return 0;
    }
    for (int f = 0; f < 2; ++f) {
        const io::fileFormat_t fmt = f ? io::_BINARY_ : io::_TEXT_;
        const std::string s = std::to_string(f);
        SphereQuarticTermTables sq;
        if (!sq.FromPseudopotentials(9, {0.0, 1.0, 0.0, 0.45, 0.0, 0.2}, mpi)) return 3;
        sq.ToFile(d + "/sq_cpp" + s, fmt, mpi);
        SphereQuarticTermTables sq2;
        sq2.FromFile(d + "/sq_cpp" + s, fmt, mpi);
        sq2.ToFile(d + "/sq_rt" + s, fmt, mpi);
        OflQuarticTermTables o;
        o.FromFile(d + "/ofl_py" + s, fmt, mpi);
        o.ToFile(d + "/ofl_cpp" + s, fmt, mpi);
        DiagonalQuadraticTermTables<dcmplx> e;
        e.FromFile(d + "/eps_py" + s, fmt, mpi);
        e.ToFile(d + "/eps_cpp" + s, fmt, mpi);
        utilities::CrsSparseMatrix<double> c;
        c.FromFile(d + "/crs_py" + s, fmt, mpi);
        c.ToFile(d + "/crs_cpp" + s, fmt, mpi);
        utilities::CrsSparseMatrix<dcmplx> cz;
        cz.FromFile(d + "/crsz_py" + s, fmt, mpi);
        cz.ToFile(d + "/crsz_cpp" + s, fmt, mpi);
        SpinlessFermionHamiltonian<double> h;
        h.EigensystemFromFile(d + "/eig_py" + s, true, true, fmt, mpi);
        h.EigensystemToFile(d + "/eig_cpp" + s, true, true, fmt, mpi);
        SpinlessFermionHamiltonian<dcmplx> hz;
        hz.EigensystemFromFile(d + "/eigz_py" + s, false, true, fmt, mpi);
        hz.EigensystemToFile(d + "/eigz_cpp" + s, true, true, fmt, mpi);
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def interop_exe(tmp_path_factory):
    if not os.path.exists(os.path.join(LIB, "libhdiag_hip.so")):
        pytest.fail("libhdiag_hip.so is not built (run __graft_entry__.build())")
    d = tmp_path_factory.mktemp("interop")
    src = d / "interop.cpp"
    src.write_text(CPP_PROGRAM)
    exe = d / "interop"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    "-I" + CPP, str(src), "-o", str(exe), "-L" + LIB, "-lhdiag_hip", "-Wl,-rpath," + LIB],
                   check=True)
    return exe


def test_cpp_and_numpy_files_byte_identical(interop_exe, tmp_path):
    rng = np.random.default_rng(7)
    kt, vt, eps = models.ofl_tables(3, 4, 8)
    crs = _random_crs(rng, 31)
    crsz = _random_crs(rng, 29, True)
    ev = np.sort(rng.normal(size=3))
    vec = rng.normal(size=(3, 40))
    evz = np.sort(rng.normal(size=2))
    vecz = rng.normal(size=(2, 15)) + 1j * rng.normal(size=(2, 15))
    for f in (0, 1):
        fileio.write_term_table(tmp_path / ("ofl_py%d" % f), "ofl", 12, vt, kt, f)
        fileio.write_term_table(tmp_path / ("eps_py%d" % f), "quadratic", 12, eps.astype(np.complex128), fmt=f)
        fileio.write_crs(tmp_path / ("crs_py%d" % f), *crs, fmt=f)
        fileio.write_crs(tmp_path / ("crsz_py%d" % f), *crsz, fmt=f)
        fileio.write_eigensystem(tmp_path / ("eig_py%d" % f), ev, vec, f)
        fileio.write_eigensystem(tmp_path / ("eigz_py%d" % f), evz, vecz, f)
    p = subprocess.run([str(interop_exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    skt = hd.sphere_ktable(9)
    svt = hd.sphere_vtable(9, [0.0, 1.0, 0.0, 0.45, 0.0, 0.2])
    for f in (0, 1):
        fileio.write_term_table(tmp_path / ("sq_py%d" % f), "sphere", 10, svt, skt, f)
        for a, b in (("sq_cpp", "sq_py"), ("sq_rt", "sq_py"), ("ofl_cpp", "ofl_py"), ("eps_cpp", "eps_py"),
                     ("crs_cpp", "crs_py"), ("crsz_cpp", "crsz_py"), ("eig_cpp", "eig_py"),
                     ("eigz_cpp", "eigz_py")):
            x = (tmp_path / ("%s%d" % (a, f))).read_bytes()
            y = (tmp_path / ("%s%d" % (b, f))).read_bytes()
            assert x == y, (a, f)


def test_cpp_reader_errors(interop_exe, tmp_path):
    # a wrong header sets the exit flag; ExitFlagTest ends the run (term_tables_base.hpp:268-275)
    (tmp_path / "bad.dat").write_bytes(struct.pack("<I", 5) + b"hello" + struct.pack("<II", 4, 3))
    p = subprocess.run([str(interop_exe), str(tmp_path), "bad-header"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "wrong file header hello" in p.stderr
    p = subprocess.run([str(interop_exe), str(tmp_path), "missing"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "Could not open file" in p.stderr


# ------------------------------------------------------------------ GPU: the driver's files

def _run(args):
    p = subprocess.run([DRIVER] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    import json
    return json.loads(p.stdout.strip().splitlines()[-1])


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [fileio.BINARY, fileio.TEXT])
def test_driver_hamiltonian_terms_and_eigensystem_files(oracle, tmp_path, fmt):
    say("drop-in files N=6 2S=15 lz=0, format", fmt)
    vm = models.coulomb_list(15)
    out = tmp_path / "out"
    out.mkdir()
    base = ["--nbr-particles", 6, "--nbr-orbitals", 16, "--lz", 0, "--nbr-eigenvalues", 3, "--tol", 1e-12,
            "--file-format", fmt, "--out-path", out, "--out-file-name", "sphere"]
    r = _run(base + ["--pseudopotentials", ",".join("%.17g" % v for v in vm), "--store-terms", 1,
                     "--hamiltonian-file", 1, "--eigenvectors-file", 1])
    dim = r["dim"]
    bas = oracle.basis(oracle.sector(oracle.SECTOR_LZ, 0, 15), 6, 16)
    kt, vt = hd.sphere_ktable(15), hd.sphere_vtable(15, vm)
    up = oracle.build(6, 16, bas, oracle.LAYOUT_SPHERE, 16, kt, vt, full=False)
    rp, col, val = fileio.read_crs(out / "sphere_sector_0_hamiltonian_0.dat", fmt)
    assert np.array_equal(rp, up.rowptr) and np.array_equal(col, up.col)
    tol = 0 if fmt == fileio.BINARY else 1e-14
    assert np.all(np.abs(val - up.values()) <= tol * np.abs(up.values()))
    k, v, t = fileio.read_term_table(out / "quartic_coefficient_table_L_16.dat", "sphere", fmt)
    assert k == 16 and np.array_equal(t, kt) and np.all(np.abs(v - vt) <= tol * np.abs(vt))
    ev, vec = fileio.read_eigensystem(out / "sphere_sector_0_eigensystem.dat", fmt)
    assert np.allclose(ev, r["eigenvalues"], rtol=1e-14, atol=0)
    M = up.to_dense(dim)
    want = np.linalg.eigvalsh(M)[:3]
    assert np.max(np.abs(ev - want)) < 1e-10
    for i in range(3):
        assert np.linalg.norm(M @ vec[i] - ev[i] * vec[i]) < 1e-9
    # the stored tables drive the same run (--retrieve-terms)
    r2 = _run(base + ["--retrieve-terms", 1])
    assert np.allclose(r2["eigenvalues"], r["eigenvalues"], rtol=1e-12, atol=1e-12)
