"""The reference program's on-disk formats, read and written with numpy.

Each function states the byte layout it follows and the reference code that
writes/reads that layout, so files move between this package and the
reference's own binaries in both directions:

* CRS matrices: ``CrsSparseMatrix::ToFile/FromFile``
  (src/utilities/data_structures/crs_sparse_matrix.hpp:662-751), which
  ``HamiltonianBase::HamiltonianToFile`` (src/hamiltonians/hamiltonian_base.hpp:1123-1153)
  uses for ``--hamiltonian-file``.
* Term tables: ``TermTables::ToFileBase/FromFileBase``
  (src/hamiltonians/term_tables_base.hpp:179-345) for ``--store-terms`` /
  ``--retrieve-terms``.
* Eigensystems: ``HamiltonianBase::EigensystemToFile/EigensystemFromFile``
  (src/hamiltonians/hamiltonian_base.hpp:1204-1432) for ``--eigenvalues-file`` /
  ``--eigenvectors-file``.
* ARPACK start vectors: the raw ``T[dim]`` file ``ArpackWrapper`` reads for
  ``--use-initial`` and writes for ``--store-final``
  (src/utilities/wrappers/arpack_wrapper.hpp:753-774, 855-870).

Text fields use 15 significant digits (``std::setprecision(15)``,
src/utilities/wrappers/io_wrapper.hpp:77-93), so text files are lossy;
binary files are exact.  All binary integers are little-endian.
"""
import numpy as np

TEXT, BINARY = 0, 1  # io::fileFormat_t (io_wrapper.hpp:24)
TERM_HEADER = "# File contains term table data in array format"  # term_tables_base.hpp:63


def _fmt(x):
    """One value as the reference's ToStream prints it: %.15g, complex as re<TAB>im."""
    if np.iscomplexobj(x):
        return "%.15g\t%.15g" % (x.real, x.imag)
    return "%.15g" % x


def _tokens(path):
    with open(path) as f:
        return f.read().split()


# ----------------------------------------------------------------- CRS

def write_crs(path, rowptr, col, val, fmt=BINARY):
    """crs_sparse_matrix.hpp:662-700.  Binary: uint32 nnz, uint32 rows, int32
    row starts (rows + 1), int32 columns, values (float64 or complex128).
    Text: nnz, rows, the first `rows` row starts, then "row col value" lines."""
    rowptr = np.asarray(rowptr, dtype=np.int64)
    col = np.asarray(col, dtype=np.int32)
    val = np.asarray(val)
    nnz, rows = len(val), len(rowptr) - 1
    if rowptr[-1] != nnz or len(col) != nnz:
        raise ValueError("inconsistent CRS arrays")
    if nnz >= 2 ** 31:
        raise ValueError("the reference's CRS row starts are int32")
    if fmt == BINARY:
        with open(path, "wb") as f:
            f.write(np.array([nnz, rows], dtype="<u4").tobytes())
            f.write(rowptr.astype("<i4").tobytes())
            f.write(col.astype("<i4").tobytes())
            f.write(val.astype("<c16" if np.iscomplexobj(val) else "<f8").tobytes())
        return
    with open(path, "w") as f:
        f.write("%d\n%d\n" % (nnz, rows))
        f.write("".join("%d\n" % r for r in rowptr[:-1]))
        for i in range(rows):
            for j in range(rowptr[i], rowptr[i + 1]):
                f.write("%d %d %s\n" % (i, col[j], _fmt(val[j])))


def read_crs(path, fmt=BINARY, complex_values=False):
    """crs_sparse_matrix.hpp:705-751.  Returns (rowptr int64, col int32, values).
    The text layout has no final row start: it is nnz (from the first line)."""
    if fmt == BINARY:
        raw = open(path, "rb").read()
        nnz, rows = (int(v) for v in np.frombuffer(raw, dtype="<u4", count=2))
        o = 8
        rowptr = np.frombuffer(raw, dtype="<i4", count=rows + 1, offset=o).astype(np.int64)
        o += 4 * (rows + 1)
        col = np.frombuffer(raw, dtype="<i4", count=nnz, offset=o).copy()
        o += 4 * nnz
        vdt = "<c16" if complex_values else "<f8"
        val = np.frombuffer(raw, dtype=vdt, count=nnz, offset=o).copy()
        if o + val.nbytes != len(raw):
            raise ValueError("%s: %d trailing bytes" % (path, len(raw) - o - val.nbytes))
        return rowptr, col, val
    t = _tokens(path)
    nnz, rows = int(t[0]), int(t[1])
    rowptr = np.array([int(x) for x in t[2:2 + rows]] + [nnz], dtype=np.int64)
    per = 4 if complex_values else 3
    body = t[2 + rows:]
    if len(body) != per * nnz:
        raise ValueError("%s: expected %d entries" % (path, nnz))
    col = np.array(body[1::per], dtype=np.int32)
    if complex_values:
        val = np.array(body[2::per], dtype=np.float64) + 1j * np.array(body[3::per], dtype=np.float64)
    else:
        val = np.array(body[2::per], dtype=np.float64)
    return rowptr, col, val


# ----------------------------------------------------------------- term tables

def _sphere_k1(kt, kmax, k2, k3, k4):
    return int(kt[(k4 * kmax + k3) * kmax + k2])


def _fold(kmax, k3, k4):
    t1, t2 = min(k3, k4), max(k3, k4)
    return t1 * kmax + t2 - t1 * (t1 + 1) // 2


def table_dim(layout, kmax):
    """CalculateDim: kMax^3 (sphere quartic), (kMax+1) kMax^2 / 2 (flux-lattice
    quartic), kMax (quadratic) (term_tables.cpp of each model)."""
    return {"sphere": kmax ** 3, "ofl": (kmax + 1) * kmax * kmax // 2, "quadratic": kmax}[layout]


def _get(layout, kt, vt, kmax, k2, k3, k4):
    """(k1, V) as GetK1 then GetVkkkk(k1, k2, k3, k4) return them
    (sphere term_tables.cpp:109-156; flux lattice term_tables.cpp:108-175)."""
    if layout == "sphere":
        i = (k4 * kmax + k3) * kmax + k2
        return int(kt[i]), vt[i]
    r = _fold(kmax, k3, k4) * kmax
    k1 = int(kt[r + k2])
    return k1, vt[r + (k1 if k4 > k3 else k2)]


def _set(layout, kt, vt, kmax, k1, k2, k3, k4, v):
    if layout == "sphere":
        i = (k4 * kmax + k3) * kmax + k2
        kt[i], vt[i] = k1, v
        return
    r = _fold(kmax, k3, k4) * kmax
    kt[r + k2] = k1
    vt[r + (k1 if k4 > k3 else k2)] = v


def write_term_table(path, layout, kmax, vtable, ktable=None, fmt=BINARY):
    """term_tables_base.hpp:179-247.  layout: "quadratic" (2 labels, E_k),
    "sphere" or "ofl" (4 labels).  Binary: uint32 header length, header, uint32
    labels, kMax in 4 bytes (the reference copies 4 bytes from its 2-byte
    m_kMax, so the upper two are whatever follows it in memory; zero here),
    then uint16 k table (4 labels) and the values.  Text: header, labels, kMax,
    then "k1 k1 V" per k or "k1 k2 k3 k4 V" over k4, k3, k2 (tab separated)."""
    labels = 2 if layout == "quadratic" else 4
    vt = np.asarray(vtable)
    n = table_dim(layout, kmax)
    if len(vt) != n or (labels == 4 and len(ktable) != n):
        raise ValueError("table size does not match kMax %d" % kmax)
    if fmt == BINARY:
        with open(path, "wb") as f:
            f.write(np.array([len(TERM_HEADER)], dtype="<u4").tobytes())
            f.write(TERM_HEADER.encode())
            f.write(np.array([labels, kmax], dtype="<u4").tobytes())
            if labels == 4:
                f.write(np.asarray(ktable).astype("<u2").tobytes())
            f.write(vt.astype("<c16" if np.iscomplexobj(vt) else "<f8").tobytes())
        return
    with open(path, "w") as f:
        f.write("%s\n%d\n%d\n" % (TERM_HEADER, labels, kmax))
        if labels == 2:
            for k in range(kmax):
                f.write("%d\t%d\t%s\n" % (k, k, _fmt(vt[k])))
            return
        kt = np.asarray(ktable)
        for k4 in range(kmax):
            for k3 in range(kmax):
                for k2 in range(kmax):
                    k1, v = _get(layout, kt, vt, kmax, k2, k3, k4)
                    f.write("%d\t%d\t%d\t%d\t%s\n" % (k1, k2, k3, k4, _fmt(v)))


def read_term_table(path, layout, fmt=BINARY, complex_values=False):
    """term_tables_base.hpp:252-345.  Returns (kmax, vtable, ktable or None).
    A wrong header raises ValueError (the reference sets its exit flag); the
    upper two bytes of the binary kMax field are ignored."""
    vdt = np.complex128 if complex_values else np.float64
    if fmt == BINARY:
        raw = open(path, "rb").read()
        hl = int(np.frombuffer(raw, dtype="<u4", count=1)[0])
        if raw[4:4 + hl].decode(errors="replace") != TERM_HEADER:
            raise ValueError("wrong file header")
        o = 4 + hl
        labels, kfield = (int(v) for v in np.frombuffer(raw, dtype="<u4", count=2, offset=o))
        kmax = kfield & 0xFFFF
        o += 8
        n = table_dim(layout, kmax)
        kt = None
        if labels == 4:
            kt = np.frombuffer(raw, dtype="<u2", count=n, offset=o).copy()
            o += 2 * n
        vt = np.frombuffer(raw, dtype="<c16" if complex_values else "<f8", count=n, offset=o).astype(vdt)
        return kmax, vt, kt
    with open(path) as f:
        if f.readline().rstrip("\n") != TERM_HEADER:
            raise ValueError("wrong file header")
        t = f.read().split()
    labels, kmax = int(t[0]), int(t[1])
    n = table_dim(layout, kmax)
    vt = np.zeros(n, dtype=vdt)
    body = t[2:]
    per = 2 + (2 if complex_values else 1) if labels == 2 else 4 + (2 if complex_values else 1)

    def value(row):
        x = body[row * per:(row + 1) * per]
        return complex(float(x[-2]), float(x[-1])) if complex_values else float(x[-1])

    if labels == 2:
        for k in range(kmax):
            vt[k] = value(k)
        return kmax, vt, None
    kt = np.zeros(n, dtype=np.uint16)
    row = 0
    for k4 in range(kmax):
        for k3 in range(kmax):
            for k2 in range(kmax):
                k1 = int(body[row * per])
                _set(layout, kt, vt, kmax, k1, k2, k3, k4, value(row))
                row += 1
    return kmax, vt, kt


# ----------------------------------------------------------------- eigensystems

def write_eigensystem(path, evals, vectors=None, fmt=BINARY):
    """hamiltonian_base.hpp:1204-1298.  Binary: uint32 nev, float64 evals, then
    (with vectors) uint32 dim and vectors[nev][dim].  Text: "##  NBR EIGENVALUES",
    nev, "##  EIGENVALUES", values, "##  FOCK SPACE DIMENSION", dim,
    "##  EIGENVECTORS", one entry per line."""
    evals = np.asarray(evals, dtype=np.float64)
    nev = len(evals)
    if vectors is not None:
        vectors = np.asarray(vectors).reshape(nev, -1)
    if fmt == BINARY:
        with open(path, "wb") as f:
            f.write(np.array([nev], dtype="<u4").tobytes())
            f.write(evals.astype("<f8").tobytes())
            if vectors is not None:
                f.write(np.array([vectors.shape[1] & 0xFFFFFFFF], dtype="<u4").tobytes())
                f.write(vectors.astype("<c16" if np.iscomplexobj(vectors) else "<f8").tobytes())
        return
    with open(path, "w") as f:
        f.write("##  NBR EIGENVALUES\n%d\n##  EIGENVALUES\n" % nev)
        f.write("".join("%.15g\n" % e for e in evals))
        if vectors is not None:
            f.write("##  FOCK SPACE DIMENSION\n%d\n##  EIGENVECTORS\n" % vectors.shape[1])
            f.write("".join(_fmt(x) + "\n" for x in vectors.ravel()))


def read_eigensystem(path, fmt=BINARY, complex_values=False, vectors=True):
    """hamiltonian_base.hpp:1304-1432.  Returns (evals, vectors[nev][dim] or None)."""
    vdt = "<c16" if complex_values else "<f8"
    if fmt == BINARY:
        raw = open(path, "rb").read()
        nev = int(np.frombuffer(raw, dtype="<u4", count=1)[0])
        evals = np.frombuffer(raw, dtype="<f8", count=nev, offset=4).copy()
        o = 4 + 8 * nev
        if not vectors or o == len(raw):
            return evals, None
        dim = int(np.frombuffer(raw, dtype="<u4", count=1, offset=o)[0])
        vec = np.frombuffer(raw, dtype=vdt, count=nev * dim, offset=o + 4).copy().reshape(nev, dim)
        return evals, vec
    lines = open(path).read().split("\n")
    i = lines.index("##  NBR EIGENVALUES")
    nev = int(lines[i + 1])
    j = lines.index("##  EIGENVALUES")
    evals = np.array([float(x) for x in lines[j + 1:j + 1 + nev]])
    if not vectors or "##  FOCK SPACE DIMENSION" not in lines:
        return evals, None
    k = lines.index("##  FOCK SPACE DIMENSION")
    dim = int(lines[k + 1])
    m = lines.index("##  EIGENVECTORS")
    body = lines[m + 1:m + 1 + nev * dim]
    if complex_values:
        vec = np.array([complex(*(float(y) for y in x.split("\t"))) for x in body])
    else:
        vec = np.array([float(x) for x in body])
    return evals, vec.reshape(nev, dim)


# ----------------------------------------------------------------- start vectors

def write_initial_vector(path, vec):
    """The raw T[dim] file ArpackWrapper reads with --use-initial
    (arpack_wrapper.hpp:753-774): float64 or complex128, no header."""
    v = np.asarray(vec)
    v.astype("<c16" if np.iscomplexobj(v) else "<f8").tofile(path)


def read_initial_vector(path, complex_values=False):
    return np.fromfile(path, dtype="<c16" if complex_values else "<f8")
