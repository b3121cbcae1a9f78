"""polatory's binary model and interpolant files (Model::save/load,
Interpolant::save/load).

The expected bytes are built here from the stream rules of
common/io.hpp:15-186 and the field orders of rbf/rbf_io.hpp:9-35,
model.hpp:304-320, geometry/bbox3d.hpp:82-98 and interpolant.hpp:345-367:
trivially copyable values as their native bytes, std::string and std::vector
as a size_t count then the elements, dynamic Eigen matrices as their dynamic
extents (Eigen::Index) then the coefficients in storage order (polatory's
Mat<M, N> is row-major for N > 1). No file written by polatory itself exists
in the reference tree, so the format is pinned by this independent encoder
and decoder only.
"""
import struct

import numpy as np
import pytest

import polatory_amd as pa


def enc_rbf(name, params, aniso):
    b = struct.pack("<Q", len(name)) + name.encode()
    b += struct.pack("<Q", len(params)) + struct.pack(f"<{len(params)}d", *params)
    b += struct.pack("<9d", *np.asarray(aniso, dtype=float).reshape(9))  # Mat3, row-major
    return b


def enc_model(rbfs, poly_degree, nugget):
    b = struct.pack("<Q", len(rbfs))
    for r in rbfs:
        b += enc_rbf(*r)
    return b + struct.pack("<i", poly_degree) + struct.pack("<d", nugget)


class Reader:
    def __init__(self, b):
        self.b, self.o = b, 0

    def take(self, fmt):
        v = struct.unpack_from(fmt, self.b, self.o)
        self.o += struct.calcsize(fmt)
        return v

    def model(self):
        (n,) = self.take("<Q")
        rbfs = []
        for _ in range(n):
            (ln,) = self.take("<Q")
            name = self.b[self.o:self.o + ln].decode()
            self.o += ln
            (np_,) = self.take("<Q")
            params = list(self.take(f"<{np_}d"))
            aniso = np.array(self.take("<9d")).reshape(3, 3)
            rbfs.append((name, params, aniso))
        (deg,) = self.take("<i")
        (nug,) = self.take("<d")
        return rbfs, deg, nug

    def points(self):
        (rows,) = self.take("<q")
        return np.array(self.take(f"<{3 * rows}d")).reshape(rows, 3)


ANISO = np.array([[1.5, 0.2, 0.0], [-0.1, 0.8, 0.3], [0.0, 0.4, 1.2]])


def test_model_file_layout(tmp_path):
    bh = pa.Biharmonic3D([2.0, 0.1])
    bh.set_anisotropy(ANISO)
    ex = pa.CovExponential([1.5, 0.7])
    m = pa.Model([bh, ex], poly_degree=1, nugget=0.25)
    f = tmp_path / "m.bin"
    m.save(f)
    b = f.read_bytes()
    assert b == enc_model([("bh3", [2.0, 0.1], ANISO), ("exp", [1.5, 0.7], np.eye(3))], 1, 0.25)
    r = Reader(b)
    rbfs, deg, nug = r.model()
    assert r.o == len(b)
    assert [x[0] for x in rbfs] == ["bh3", "exp"]
    assert np.array_equal(rbfs[0][2], ANISO)  # a transposed (column-major) write fails here
    assert (deg, nug) == (1, 0.25)


def test_model_load_from_reference_layout(tmp_path):
    f = tmp_path / "m.bin"
    f.write_bytes(enc_model([("th3", [0.5, 0.0], ANISO)], 2, 1e-3))
    m = pa.Model.load(f)
    assert len(m.rbfs) == 1 and isinstance(m.rbfs[0], pa.Triharmonic3D)
    assert m.rbfs[0].parameters() == [0.5, 0.0]
    assert np.array_equal(m.rbfs[0].anisotropy(), ANISO)
    assert m.poly_degree() == 2 and m.nugget() == 1e-3
    # round trip is byte-identical
    g = tmp_path / "m2.bin"
    m.save(g)
    assert g.read_bytes() == f.read_bytes()


def test_model_load_errors(tmp_path):
    f = tmp_path / "bad.bin"
    f.write_bytes(enc_model([("xyz", [1.0, 0.0], np.eye(3))], 1, 0.0))
    with pytest.raises(RuntimeError, match="unknown RBF name"):
        pa.Model.load(f)
    f.write_bytes(enc_model([("gau", [1.0, 1.0], np.eye(3))], 1, 0.0))
    with pytest.raises(NotImplementedError):
        pa.Model.load(f)
    f.write_bytes(enc_model([("bh3", [1.0, 0.0, 3.0], np.eye(3))], 1, 0.0))
    with pytest.raises(ValueError, match="params.size"):
        pa.Model.load(f)
    good = enc_model([("bh3", [1.0, 0.0], np.eye(3))], 1, 0.0)
    for cut in (3, 20, len(good) - 1):
        f.write_bytes(good[:cut])
        with pytest.raises(RuntimeError, match="end of file"):
            pa.Model.load(f)
    f.write_bytes(struct.pack("<Q", 1 << 40))
    with pytest.raises(RuntimeError, match="corrupt"):
        pa.Model.load(f)
    with pytest.raises(RuntimeError, match="cannot open"):
        pa.Model.load(tmp_path / "missing.bin")


def enc_interpolant(model_bytes, fitted, centers, grad_centers, lo, hi, weights):
    b = model_bytes + struct.pack("<?", fitted)
    for p in (centers, grad_centers):
        p = np.asarray(p, dtype=float).reshape(-1, 3)
        b += struct.pack("<q", p.shape[0]) + p.astype("<f8").tobytes()
    b += struct.pack("<3d", *lo) + struct.pack("<3d", *hi)
    w = np.asarray(weights, dtype=float)
    return b + struct.pack("<q", w.shape[0]) + w.astype("<f8").tobytes()


@pytest.mark.gpu
def test_interpolant_file_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 1, (2000, 3))
    vals = np.sin(4 * pts[:, 0]) + pts[:, 1] * pts[:, 2]
    rbf = pa.Biharmonic3D([1.0, 0.0])
    rbf.set_anisotropy(ANISO)
    ip = pa.Interpolant(pa.Model(rbf, poly_degree=1))
    ip.fit(pts, vals, 1e-6)
    f = tmp_path / "ip.bin"
    ip.save(f)
    b = f.read_bytes()
    w = ip.weights()
    expect = enc_interpolant(enc_model([("bh3", [1.0, 0.0], ANISO)], 1, 0.0), True, pts,
                             np.zeros((0, 3)), pts.min(axis=0), pts.max(axis=0), w)
    assert b == expect
    ip2 = pa.Interpolant.load(f)
    assert ip2.fitted()
    assert np.array_equal(ip2.centers(), pts) and np.array_equal(ip2.weights(), w)
    q = rng.uniform(0, 1, (500, 3))
    assert np.array_equal(ip2.evaluate(q), ip.evaluate(q))
    # unfitted: empty centres and weights, the default bbox (+inf, -inf)
    u = pa.Interpolant(pa.Model(pa.CovExponential([1.0, 0.5]), poly_degree=0))
    u.save(f)
    assert f.read_bytes() == enc_interpolant(enc_model([("exp", [1.0, 0.5], np.eye(3))], 0, 0.0), False,
                                             np.zeros((0, 3)), np.zeros((0, 3)), [np.inf] * 3,
                                             [-np.inf] * 3, [])
    assert not pa.Interpolant.load(f).fitted()
    # with gradient centres (interpolant.hpp:340-367): grad_centers_ after the
    # centres, the box of both, weights [w; g; c]
    gp = rng.uniform(0, 1, (300, 3))
    # f = 4 x + y z: values at 500 points, gradients (4, z, y) at 300 more
    gvals = np.concatenate([4 * pts[:500, 0] + pts[:500, 1] * pts[:500, 2],
                            np.stack([np.full(300, 4.0), gp[:, 2], gp[:, 1]], axis=1).ravel()])
    hip_ = pa.Interpolant(pa.Model(pa.Triharmonic3D([1.0]), poly_degree=1))
    hip_.fit_hermite(pts[:500], gp, gvals, 1e-5, 1e-5)
    hip_.save(f)
    hw = hip_.weights()
    assert hw.shape[0] == 500 + 900 + 4
    allc = np.vstack([pts[:500], gp])
    assert f.read_bytes() == enc_interpolant(enc_model([("th3", [1.0, 0.0], np.eye(3))], 1, 0.0), True, pts[:500],
                                             gp, allc.min(axis=0), allc.max(axis=0), hw)
    h2 = pa.Interpolant.load(f)
    assert np.array_equal(h2.grad_centers(), gp) and np.array_equal(h2.weights(), hw)
    assert np.array_equal(h2.evaluate_hermite(q, q[:50]), hip_.evaluate_hermite(q, q[:50]))
