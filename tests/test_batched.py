"""The batched form (sif_amd_exodus_batch): boxes of uneven size, each on
the grid matching its particle lattice, run side by side; every box's
catalogue must be the bits of that box run alone (CIC, contrast, mesh,
sif_amd_exodus_device), whatever the tuning (boxes side by side, CIC tile,
growth split, growth batch window), and the autotuner's choice too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THRESH = -0.8


@pytest.fixture(scope="module")
def boxes(native):
    from sif_amd import device as D
    from sif_amd import zeldovich
    rng = np.random.default_rng(5)
    out = []
    for b, n_side in enumerate([24, 56, 32, 40, 48, 24, 64, 40, 32, 48]):
        box = n_side * 100.0 / 64.0
        x, y, z = zeldovich.generate(n_side, box, 700 + b)
        out.append(dict(d=(D.to_device(x), D.to_device(y), D.to_device(z)), n=len(x),
                        n_cells=n_side, mesh_cells=max(4, int(round(box / 10.0))), box=box,
                        radii=zeldovich.ladder(3.0, min(20.0, box / 4.0), 0.05)))
        assert rng is not None
    return out


def _plain(D, bx):
    dx, dy, dz = bx["d"]
    g = D.cic(dx, dy, dz, bx["n"], bx["n_cells"], bx["box"])
    D.density_contrast(g, bx["n_cells"] ** 3)
    m = D.Mesh(dx, dy, dz, bx["n"], bx["mesh_cells"], bx["box"], indices=False)
    cat, _, _ = D.exodus(g, bx["n_cells"], bx["box"], m, bx["radii"], THRESH, restore=False)
    return cat


def test_batch_equals_each_box_alone_under_two_tunings(boxes):
    from sif_amd import device as D
    plain = [_plain(D, bx) for bx in boxes]
    assert sum(len(c) for c in plain) > 50
    a = D.exodus_batch(boxes, THRESH, tuning=dict(workers=1, cic_tile=8, split_waves=4096))
    b = D.exodus_batch(boxes, THRESH, tuning=dict(workers=4, cic_tile=16, split_waves=512,
                                                  batch_floor=16, batch_ceiling=64))
    for k, (p, ca, cb) in enumerate(zip(plain, a, b)):
        np.testing.assert_array_equal(ca.view(np.uint32), p.view(np.uint32), err_msg=f"box {k}")
        np.testing.assert_array_equal(cb.view(np.uint32), p.view(np.uint32), err_msg=f"box {k}")


def test_autotuned_batch(boxes):
    from sif_amd import device as D
    t = D.batch_autotune(boxes, THRESH, probe=4)
    assert t["workers"] >= 1 and t["cic_tile"] in (8, 16) and t["split_waves"] >= 64
    cats = D.exodus_batch(boxes, THRESH, tuning=t)
    ref = D.exodus_batch(boxes, THRESH, tuning=dict(workers=1))
    for c, r in zip(cats, ref):
        np.testing.assert_array_equal(c.view(np.uint32), r.view(np.uint32))


def test_batch_refuses_and_names_the_box(boxes):
    from sif_amd import device as D
    bad = [dict(bx) for bx in boxes[:3]]
    bad[1]["radii"] = np.array([5.0, -1.0], np.float32)
    with pytest.raises(ValueError, match="radii"):
        D.exodus_batch(bad, THRESH)
