"""OBJ and solve-record files (io/obj_io.cpp:12-102 semantics)."""

import numpy as np

from chysx import io, scenes


def test_obj_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(0)
    p, t = scenes.grid(7)
    p = p + rng.normal(size=p.shape).astype(np.float32) * 1e-3
    io.save_obj(str(tmp_path / "m.obj"), p, t)
    p2, t2 = io.load_obj(str(tmp_path / "m.obj"))
    np.testing.assert_array_equal(p2, p.astype(np.float32))
    np.testing.assert_array_equal(t2, t)
    head = open(tmp_path / "m.obj").readline()
    assert head.startswith("# ChysX OBJ export") and "verts=49" in head


def test_obj_reads_slash_faces_and_skips_other_lines(tmp_path):
    (tmp_path / "q.obj").write_text(
        "# comment\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvt 0 0\nvn 0 0 1\n"
        "f 1/1/1 2/2/1 3/3/1\nf 2//1 4//1 3//1\ng group\n")
    p, t = io.load_obj(str(tmp_path / "q.obj"))
    assert p.shape == (4, 3)
    np.testing.assert_array_equal(t, [[0, 1, 2], [1, 3, 2]])


def test_solve_record_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    dump = (rng.normal(size=(5, 3, 3)).astype(np.float32), np.array([0, 1, 2, 2, 3, 4], np.int32),
            np.array([1, 0, 4, 3], np.int32), rng.normal(size=(4, 3, 3)).astype(np.float32),
            rng.normal(size=(5, 3)).astype(np.float32), rng.normal(size=(5, 3)).astype(np.float32))
    io.save_solve(str(tmp_path / "s.npz"), dump)
    back = io.load_solve(str(tmp_path / "s.npz"))
    for a, b in zip(dump, back):
        np.testing.assert_array_equal(a, b)
        assert a.dtype == b.dtype
