"""The PLY reader (tpd_load_ply, GaussianPoint::fromModel of GaussianGeometry.cpp:59-127) on the
CPU: bit-exact against the reference's own fromModel compiled from its sources (oracle/_ref, with
its vendored miniply) where that is built, and against the oracle's C restatement of the
per-vertex conversion; closed forms pin sigmoid(0) = 1/2, exp(0) = 1 and the quaternion order.
Files are written here in every format the reader accepts."""
import os

import numpy as np
import pytest

REST_ORDER = [f"f_rest_{k}" for k in range(45)]


def write_ply(path, cols, fmt="binary_little_endian", pre_faces=0, dtype="float"):
    """cols: ordered dict name -> (N,) values; all vertex properties of one type."""
    names = list(cols)
    n = len(cols[names[0]])
    npt = {"float": "f4", "double": "f8"}[dtype]
    head = ["ply", f"format {fmt} 1.0", "comment written by test_ply"]
    if pre_faces:
        head += [f"element face {pre_faces}", "property list uchar int vertex_indices"]
    head += [f"element vertex {n}"] + [f"property {dtype} {k}" for k in names] + ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        if fmt == "ascii":
            for i in range(pre_faces):
                f.write(b"3 0 1 2\n")
            for i in range(n):
                f.write((" ".join(repr(float(cols[k][i])) for k in names) + "\n").encode())
            return
        end = "<" if fmt == "binary_little_endian" else ">"
        for i in range(pre_faces):
            f.write(np.array([3], np.uint8).tobytes())
            f.write(np.array([0, 1, 2], end + "i4").tobytes())
        rec = np.zeros(n, dtype=[(k, end + npt) for k in names])
        for k in names:
            rec[k] = cols[k]
        f.write(rec.tobytes())


def columns(n, seed, n_rest=45, normals=True):
    rng = np.random.default_rng(seed)
    c = {}
    for k in ("x", "y", "z"):
        c[k] = rng.normal(size=n).astype(np.float32)
    if normals:
        for k in ("nx", "ny", "nz"):
            c[k] = np.zeros(n, np.float32)
    for k in ("f_dc_0", "f_dc_1", "f_dc_2"):
        c[k] = rng.normal(size=n).astype(np.float32)
    for k in REST_ORDER[:n_rest]:
        c[k] = (0.3 * rng.normal(size=n)).astype(np.float32)
    c["opacity"] = (2.0 * rng.normal(size=n)).astype(np.float32)
    for k in ("scale_0", "scale_1", "scale_2"):
        c[k] = (rng.normal(size=n) - 4.0).astype(np.float32)
    for k in ("rot_0", "rot_1", "rot_2", "rot_3"):
        c[k] = rng.normal(size=n).astype(np.float32)
    return c


def oracle_records(orc, c, n_rest):
    n = len(c["x"])
    xyz = np.stack([c["x"], c["y"], c["z"]], 1)
    rot = np.stack([c["rot_0"], c["rot_1"], c["rot_2"], c["rot_3"]], 1)
    ls = np.stack([c["scale_0"], c["scale_1"], c["scale_2"]], 1)
    ft = np.stack([c[k] for k in ["f_dc_0", "f_dc_1", "f_dc_2"] + REST_ORDER[:n_rest]], 1)
    return orc.gaussians_from_ply_columns(xyz, c["opacity"].astype(np.float32), rot, ls, ft)


def check_file(orc, path, c, n_rest, with_reference=True):
    from torpedo_amd.engine import from_model
    got = from_model(path)
    assert got.shape == (len(c["x"]), 60)
    ora = oracle_records(orc, {k: np.asarray(v, np.float32) for k, v in c.items()}, n_rest)
    assert np.array_equal(got.view(np.uint32), ora.view(np.uint32))
    if with_reference and orc.ref_ply() is not None:
        ref = orc.ref_from_model(str(path))
        assert ref is not None, "the reference loader rejected the file"
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    return got


@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian", "ascii"])
@pytest.mark.parametrize("n_rest", [45, 9, 0])
def test_formats_and_sh_degrees(orc, tmp_path, fmt, n_rest):
    c = columns(300 if fmt != "ascii" else 40, seed=len(fmt) + n_rest, n_rest=n_rest)
    p = tmp_path / "scene.ply"
    write_ply(p, c, fmt=fmt)
    check_file(orc, p, c, n_rest)


def test_double_properties(orc, tmp_path):
    c = columns(120, seed=3)
    p = tmp_path / "d.ply"
    write_ply(p, c, fmt="binary_little_endian", dtype="double")
    check_file(orc, p, {k: np.asarray(v, np.float64).astype(np.float32) for k, v in c.items()}, 45)


def test_elements_before_the_vertices(orc, tmp_path):
    # The reference's read loop (GaussianGeometry.cpp:92-101) never advances past an element that
    # is not "vertex", so it does not return on such files; this reader skips them.  Checked
    # against the oracle's conversion only.
    c = columns(60, seed=4)
    for fmt in ("binary_little_endian", "ascii"):
        p = tmp_path / f"f_{fmt}.ply"
        write_ply(p, c, fmt=fmt, pre_faces=7)
        check_file(orc, p, c, 45, with_reference=False)


def test_closed_forms(orc, tmp_path):
    c = columns(3, seed=9, n_rest=0)
    c["opacity"][:] = 0.0
    c["scale_0"][:] = 0.0
    c["rot_0"][:] = [2.0, 0.0, 0.0]  # w
    c["rot_1"][:] = [0.0, 3.0, 0.0]  # x
    c["rot_2"][:] = [0.0, 0.0, 0.5]  # y
    c["rot_3"][:] = [0.0, 4.0, 0.0]  # z
    p = tmp_path / "c.ply"
    write_ply(p, c)
    g = check_file(orc, p, c, 0)
    assert np.all(g[:, 3] == 0.5) and np.all(g[:, 8] == 1.0) and np.all(g[:, 11] == 1.0)
    assert np.array_equal(g[0, 4:8], [0, 0, 0, 1])
    assert np.allclose(g[1, 4:8], [0.6, 0, 0.8, 0], atol=1e-7)
    assert np.array_equal(g[2, 4:8], [0, 1, 0, 0])
    assert not g[:, 15:60].any()


def test_errors(tmp_path):
    from torpedo_amd._capi import TpdError
    from torpedo_amd.engine import from_model
    with pytest.raises(TpdError):
        from_model(tmp_path / "missing.ply")
    c = columns(10, seed=1)
    del c["opacity"]
    p = tmp_path / "noop.ply"
    write_ply(p, c)
    with pytest.raises(TpdError, match="opacity"):
        from_model(p)
    c = columns(10, seed=2)
    p = tmp_path / "trunc.ply"
    write_ply(p, c)
    data = open(p, "rb").read()
    open(p, "wb").write(data[:-7])
    with pytest.raises(TpdError, match="truncated"):
        from_model(p)
    p = tmp_path / "junk.ply"
    open(p, "wb").write(b"not a ply\n")
    with pytest.raises(TpdError):
        from_model(p)


@pytest.mark.parametrize("fmt", ["binary_little_endian", "ascii"])
def test_vertex_count_beyond_the_data(tmp_path, fmt):
    """A header may claim any count: 4e9 vertices over a few bytes of data is refused before any
    buffer is sized by it, and a count above 2^32 - 1 too."""
    from torpedo_amd._capi import TpdError
    from torpedo_amd.engine import from_model
    props = ["x", "y", "z", "opacity", "rot_0", "rot_1", "rot_2", "rot_3", "scale_0", "scale_1", "scale_2",
             "f_dc_0", "f_dc_1", "f_dc_2"]
    for count, code in ((4_000_000_000, "INVALID"), (5_000_000_000, "CAPACITY")):
        head = f"ply\nformat {fmt} 1.0\nelement vertex {count}\n"
        head += "".join(f"property float {p}\n" for p in props) + "end_header\n"
        body = b"\0" * (4 * len(props) * 3) if fmt != "ascii" else b"0 " * len(props) * 3
        p = tmp_path / f"huge_{count}.ply"
        open(p, "wb").write(head.encode() + body)
        with pytest.raises(TpdError, match=code):
            from_model(p)
