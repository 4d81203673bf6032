"""Grid files in the reference's format (raw FP32 .ba1 + text .dims) and its atom-type
table (src/G3D.ml:14-64, src/lds.ml:471-555, :1600-1606, src/mol.ml:280-304)."""
import numpy as np
import pytest

from mmo_amd import gridio


def test_reduce_precision_matches_reference_rule():
    # trunc(x * 100 +- 0.5) / 100, src/utls.ml:129-132
    assert gridio.reduce_precision(0.123) == 0.12
    assert gridio.reduce_precision(0.125) == 0.13
    assert gridio.reduce_precision(-0.125) == -0.13
    assert gridio.reduce_precision(-0.004) == 0.0
    assert gridio.reduce_precision(0.0) == 0.0


def test_type_table_first_appearance_order():
    l1 = {"anum": [6, 1, 6, 8], "q": [0.1, 0.05, 0.1, -0.4]}
    l2 = {"anum": [8, 7, 6], "q": [-0.4, -0.3, 0.2]}
    types, ids = gridio.type_table([l1, l2])
    assert types == [(6, 0.1), (1, 0.05), (8, -0.4), (7, -0.3), (6, 0.2)]
    assert ids[0].tolist() == [0, 1, 0, 2] and ids[1].tolist() == [2, 3, 4]
    l3 = {"anum": [6, 6], "q": [0.101, 0.0999]}
    types, ids = gridio.type_table([l3], less_charges=True)
    assert types == [(6, 0.1)] and ids[0].tolist() == [0, 0]


def test_atom_types_file_format(tmp_path):
    fn = str(tmp_path / "atom_types.txt")
    types = [(6, 0.1), (1, 0.05), (8, -0.4123456)]
    gridio.write_atom_types(fn, types)
    assert open(fn).read() == "0 (6, 0.1)\n1 (1, 0.05)\n2 (8, -0.412346)\n"
    assert gridio.read_atom_types(fn) == [(6, 0.1), (1, 0.05), (8, -0.412346)]


def test_ba1_dims_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    grids = rng.normal(0, 100, (3, 5, 6, 7)).astype(np.float32)   # (types, z, y, x)
    base = str(tmp_path / "ligs.mol2")
    gridio.save_grids(base, grids, 0.375)
    assert open(base + ".t1.ba1.dims").read() == "step: 0.375\nx_dim: 7\ny_dim: 6\nz_dim: 5\n"
    raw = np.fromfile(base + ".t2.ba1", dtype=np.float32)
    assert raw.size == 7 * 6 * 5
    # x fastest: element (k, j, i) at i + j * x_dim + k * x_dim * y_dim
    assert raw[3 + 2 * 7 + 4 * 42] == grids[2, 4, 2, 3]
    back, step = gridio.load_grids(base, 3)
    assert step == 0.375 and back.dtype == np.float32
    assert np.array_equal(back, grids)


def test_compressed_grid_files(tmp_path):
    """The reference's compressed cache (<file>.ba1.zst, raw file removed, .dims kept):
    a standard zstd frame whose content is the raw .ba1 bytes; load_grids reads it."""
    from mmo_amd import zstd
    rng = np.random.default_rng(1)
    grids = rng.normal(0, 100, (2, 4, 5, 6)).astype(np.float32)
    grids[:, :2] = 0.0                                   # the sparse parts compress
    base = str(tmp_path / "ligs.mol2")
    gridio.save_grids(base, grids, 0.5, compress=True)
    import os
    assert not os.path.exists(base + ".t0.ba1") and os.path.exists(base + ".t0.ba1.zst")
    assert open(base + ".t0.ba1.dims").read() == "step: 0.5\nx_dim: 6\ny_dim: 5\nz_dim: 4\n"
    blob = open(base + ".t1.ba1.zst", "rb").read()
    assert blob[:4] == b"\x28\xb5\x2f\xfd"               # zstd frame magic 0xFD2FB528
    assert zstd.decompress(blob) == grids[1].tobytes()   # the raw file's bytes
    back, step = gridio.load_grids(base, 2)
    assert step == 0.5 and np.array_equal(back, grids)
    # a raw file next to the compressed one is preferred (the reference's --no-compress)
    grids[0, 3, 4, 5] = 7.0
    gridio.write_ba1(base + ".t0.ba1", grids[0], 0.5)
    back, _ = gridio.load_grids(base, 2)
    assert back[0, 3, 4, 5] == 7.0
    with pytest.raises(ValueError):
        zstd.decompress(blob, size=len(grids[1].tobytes()) + 4)
    with pytest.raises(ValueError):
        zstd.decompress(b"not a frame")


def test_roi_bild_file(tmp_path):
    """ROI.from_bild on the reference's data/ROI.bild (a copy in tests/golden), the
    writer's three lines, and the one-sphere rule."""
    import os
    from mmo_amd import roi
    here = os.path.dirname(os.path.abspath(__file__))
    c, r = roi.from_bild(os.path.join(here, "golden", "ROI.bild"))
    assert np.array_equal(c, [11.47, -3.64, -32.13]) and r == 10.0
    fn = str(tmp_path / "roi.bild")
    roi.to_bild(fn, [1.5, -2.25, 1e-3], 6.5)
    assert open(fn).read() == ".transparency 0.8\n.color blue\n.sphere 1.5 -2.25 0.001 6.5\n"
    c2, r2 = roi.from_bild(fn)
    assert np.array_equal(c2, [1.5, -2.25, 1e-3]) and r2 == 6.5
    assert roi.is_inside(c2, r2, c2 + [6.4, 0, 0]) and not roi.is_inside(c2, r2, c2 + [6.5, 0, 0])
    with open(fn, "a") as f:
        f.write(".sphere 0 0 0 1\n")
    with pytest.raises(ValueError):
        roi.from_bild(fn)
    with open(fn, "w") as f:
        f.write(".color blue\n")
    with pytest.raises(ValueError):
        roi.from_bild(fn)


def test_ba1_size_mismatch_is_an_error(tmp_path):
    fn = str(tmp_path / "x.t0.ba1")
    np.zeros(10, dtype=np.float32).tofile(fn)
    gridio.write_dims(fn + ".dims", 0.5, (2, 2, 2))
    with pytest.raises(ValueError):
        gridio.read_ba1(fn)


def test_vdw_mask_restatement_closed_form():
    """One atom of radius 1 A at a node, step 0.5 A: the nodes strictly inside are the
    27 lattice points with i^2 + j^2 + k^2 < 4 (1 + 6 + 12 + 8)."""
    from oracle import scan_ref
    m = scan_ref.vdw_mask([[2.0, 2.0, 2.0]], [1.0], 0.5, (9, 9, 9))
    assert m.sum() == 27
    assert m[4, 4, 4] and m[3, 3, 3] and not m[2, 4, 4]
    w = scan_ref.mask_words(m)
    v = 4 + 4 * 9 + 4 * 81
    assert (w[v // 32] >> (v % 32)) & 1 == 1
    assert sum(bin(int(x)).count("1") for x in w) == 27
    # 8-corner tests: a point inside the sphere's cell vs one far away
    assert scan_ref.clash_or(m, 0.5, (0, 0, 0), [2.1, 2.1, 2.1])
    assert scan_ref.clash_and(m, 0.5, (0, 0, 0), [2.1, 2.1, 2.1])
    assert not scan_ref.clash_or(m, 0.5, (0, 0, 0), [0.1, 0.1, 0.1])
    assert not scan_ref.clash_or(m, 0.5, (0, 0, 0), [-5.0, 0.1, 0.1])   # outside: free


def test_topk_restatement_order():
    from oracle import scan_ref
    e = np.array([3.0, -1.0, np.nan, -1.0, np.inf, 0.5, -2.0])
    v, i = scan_ref.topk(e, 5)
    assert i.tolist() == [6, 1, 3, 5, 0]
    v, i = scan_ref.topk(e, 7)
    assert i.tolist() == [6, 1, 3, 5, 0, 4, 2]
