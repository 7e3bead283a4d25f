"""Solution files in HDF5 format (channel/channel.c:791-1028), written and
read by the library without libhdf5 (runs without a GPU).

* The reader against a file that libhdf5 wrote: scipy's MATLAB 7.3 test file
  testhdf5_7.4_GLNX86.mat (an HDF5 file behind a 512-byte userblock) holds
  the dataset testdouble = pi/4 * (0..8) as a 9x1 float64 array, with a
  version-2 layout message, a fill-value message, a modification time and an
  attribute in its object header.
* The writer: a save_solution file of a 16x33x12 grid read back dataset by
  dataset (shapes, values, data_Ny = 2 qpts / 3), and walked by an
  independent reader below that checks the structure libhdf5 expects:
  superblock 0 with the default B-tree ranks, the root group's cached
  symbol-table addresses, the local heap's free list (1 = none, libhdf5's
  mark), full-size B-tree and symbol-table nodes, sorted names, and each
  dataset's object header (dataspace, the {real, imaginary} compound or the
  scalar type, contiguous layout at the data).
* Errors: a missing dataset, a file that is not HDF5, too small an output.
"""
import os
import struct

import numpy as np
import pytest

MAT = "/usr/local/lib/python3.10/dist-packages/scipy/io/matlab/tests/data/testhdf5_7.4_GLNX86.mat"


@pytest.fixture(scope="module")
def cc():
    import c_channel
    return c_channel


@pytest.mark.skipif(not os.path.exists(MAT), reason="scipy's MATLAB 7.3 test file is absent")
def test_reads_a_file_written_by_libhdf5(cc):
    a = cc.h5_read(MAT, "testdouble")
    assert a.shape == (9, 1) and a.dtype == np.float64
    np.testing.assert_array_equal(a[:, 0], np.pi / 4 * np.arange(9))


def _solution(tmp_path, Nx=16, Ny=33, Nz=12):
    rng = np.random.default_rng(3)
    shape = (Nz, 2, Ny - 2, Nx // 2)
    C = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    IC = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    path = str(tmp_path / "sol.hd5")
    cc_mod = __import__("c_channel")
    cc_mod.h5_write_solution(path, Nx, Ny, Nz, 0.005, 2000.0, C, IC)
    return path, C, IC


def test_solution_file_round_trip(cc, tmp_path):
    Nx, Ny, Nz = 16, 33, 12
    path, C, IC = _solution(tmp_path, Nx, Ny, Nz)
    for name, src, v in (("data_alpha", C, 0), ("data_beta", C, 1), ("data_ialpha", IC, 0),
                         ("data_ibeta", IC, 1)):
        a = cc.h5_read(path, "/" + name)
        assert a.shape == (Ny - 2, Nz, Nx // 2) and a.dtype == np.complex128
        np.testing.assert_array_equal(a, np.transpose(src[:, v], (1, 0, 2)))
    qpts = 3 * Ny // 2
    assert cc.h5_read(path, "data_Nx").tolist() == [Nx]
    assert cc.h5_read(path, "data_Ny").tolist() == [qpts * 2 // 3]
    assert cc.h5_read(path, "data_Nz").tolist() == [Nz]
    assert cc.h5_read(path, "data_dt").tolist() == [0.005]
    assert cc.h5_read(path, "data_Re").tolist() == [1 / 2000.0]      # the reference's re = 1/Re
    assert cc.h5_read(path, "data_mpg").tolist() == [0.0]


class _Walk:
    """A second, independent reading of the structure (8-byte offsets)."""

    def __init__(self, path):
        self.f = open(path, "rb").read()

    def u(self, at, n):
        return int.from_bytes(self.f[at:at + n], "little")

    def messages(self, oh):
        f = self.f
        assert f[oh] == 1, "object header version 1"
        n, size = self.u(oh + 2, 2), self.u(oh + 8, 4)
        p, out = oh + 16, []
        while p < oh + 16 + size:
            t, s = self.u(p, 2), self.u(p + 2, 2)
            assert s % 8 == 0
            out.append((t, p + 8, s))
            p += 8 + s
        assert p == oh + 16 + size and len(out) == n
        return out


def test_solution_file_structure(tmp_path):
    path, C, IC = _solution(tmp_path)
    w = _Walk(path)
    f = w.f
    assert f[:8] == b"\x89HDF\r\n\x1a\n"
    assert f[8:16] == bytes([0, 0, 0, 0, 0, 8, 8, 0])          # versions, offset/length sizes
    assert w.u(16, 2) == 4 and w.u(18, 2) == 16                  # libhdf5's default ranks
    assert w.u(24, 8) == 0 and w.u(32, 8) == 2 ** 64 - 1         # base, no free-space info
    assert w.u(40, 8) == len(f)                                  # end of file
    root_oh, cache = w.u(64, 8), w.u(72, 4)
    btree, heap = w.u(80, 8), w.u(88, 8)
    assert cache == 1
    msgs = w.messages(root_oh)
    assert [(t, w.u(p, 8), w.u(p + 8, 8)) for t, p, s in msgs if t == 0x11] == [(0x11, btree, heap)]
    assert f[heap:heap + 4] == b"HEAP" and f[heap + 4] == 0
    hsize, free, hdata = w.u(heap + 8, 8), w.u(heap + 16, 8), w.u(heap + 24, 8)
    assert free == 1 or free < hsize
    assert f[hdata:hdata + 8] == bytes(8)                       # the empty name at offset 0

    def name(off):
        return f[hdata + off:f.index(b"\0", hdata + off)].decode()
    assert f[btree:btree + 6] == b"TREE\0\0"
    used = w.u(btree + 6, 2)
    node_end = btree + 24 + 33 * 8 + 32 * 8                      # 2K+1 keys, 2K children
    assert node_end <= min(w.u(btree + 32 + 16 * i, 8) for i in range(used))
    names, ohs = [], {}
    for i in range(used):
        snod = w.u(btree + 32 + 16 * i, 8)
        assert f[snod:snod + 6] == b"SNOD\x01\0"
        for k in range(w.u(snod + 6, 2)):
            e = snod + 8 + 40 * k
            nm = name(w.u(e, 8))
            names.append(nm)
            ohs[nm] = w.u(e + 8, 8)
        assert name(w.u(btree + 40 + 16 * i, 8)) == names[-1]    # key after child i
    assert names == sorted(names, key=lambda s: s.encode())
    assert set(names) == {"data_alpha", "data_beta", "data_ialpha", "data_ibeta", "data_Nx",
                          "data_Ny", "data_Nz", "data_dt", "data_Re", "data_mpg"}
    for nm, oh in ohs.items():
        m = {t: (p, s) for t, p, s in w.messages(oh)}
        assert {1, 3, 5, 8} <= set(m)
        p, s = m[1]
        rank = f[p + 1]
        dims = [w.u(p + 8 + 8 * i, 8) for i in range(rank)]
        p, s = m[3]
        cls, size = f[p] & 15, w.u(p + 4, 4)
        if nm in ("data_alpha", "data_beta", "data_ialpha", "data_ibeta"):
            assert dims == [31, 12, 8] and cls == 6 and size == 16 and w.u(p + 1, 2) == 2
            assert f[p + 8:p + 13] == b"real\0" and w.u(p + 16, 4) == 0
            q = p + 16 + 32                                      # first member's type
            assert f[q] & 15 == 1 and w.u(q + 4, 4) == 8
            q += 20
            assert f[q:q + 10] == b"imaginary\0" and w.u(q + 16, 4) == 8
        else:
            assert dims == [1] and cls == (0 if nm[-2:] in ("Nx", "Ny", "Nz") else 1)
        p, s = m[8]
        assert f[p] == 3 and f[p + 1] == 1                       # layout 3, contiguous
        addr, n = w.u(p + 2, 8), w.u(p + 10, 8)
        assert n == size * int(np.prod(dims)) and addr + n <= len(f)
        if nm == "data_alpha":
            raw = np.frombuffer(f[addr:addr + n], np.complex128).reshape(dims)
            np.testing.assert_array_equal(raw, np.transpose(C[:, 0], (1, 0, 2)))
        if nm == "data_Re":
            assert struct.unpack("<d", f[addr:addr + 8])[0] == 1 / 2000.0


def test_errors(cc, tmp_path):
    path, C, IC = _solution(tmp_path)
    with pytest.raises(RuntimeError, match="no dataset"):
        cc.h5_read(path, "data_gamma")
    bad = tmp_path / "not.hd5"
    bad.write_bytes(b"plain text, not HDF5" * 40)
    with pytest.raises(RuntimeError, match="not an HDF5 file"):
        cc.h5_read(str(bad), "data_alpha")
    import ctypes
    out = np.zeros(4)
    assert cc._L.sc_h5_read_doubles(path.encode(), b"data_alpha", out.ctypes.data, 4) == -3
    with pytest.raises(TypeError):
        cc.h5_write_solution(path, 16, 33, 12, 0.005, 2000.0, C[:, :, :3], IC)
