"""Argument checks of the Python API that run before any C call (no GPU needed): inputs reach the kernels as raw
device pointers, so host data, wrong shapes, wrong dtypes and mismatched lengths must be refused up front."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from blazert_amd import Scene  # noqa: E402
from blazert_amd import scene as S  # noqa: E402


def test_host_arrays_are_refused():
    with pytest.raises(ValueError, match="CUDA"):
        S._columns(np.zeros((4, 3)))
    with pytest.raises(ValueError, match="CUDA"):
        S._columns(torch.zeros((4, 3), dtype=torch.float64))
    with pytest.raises(ValueError, match="CUDA"):
        S._columns((torch.zeros(4), torch.zeros(4), torch.zeros(4)))


def test_shapes_dtypes_and_lengths_are_refused():
    with pytest.raises(ValueError, match=r"\(n,3\)"):
        S._columns(torch.zeros((4, 2)))
    with pytest.raises(ValueError, match="floating"):
        S._columns(torch.zeros((4, 3), dtype=torch.int32))
    with pytest.raises(ValueError, match="differ in length"):
        S._columns((torch.zeros(4), torch.zeros(4), torch.zeros(5)))
    with pytest.raises(ValueError, match="1-D"):
        S._columns((torch.zeros(4, 1), torch.zeros(4), torch.zeros(4)))


def test_tmin_tmax_columns():
    with pytest.raises(ValueError, match="scalar or a CUDA tensor"):
        S._scalar_col(np.zeros(4), 4, "cpu")
    with pytest.raises(ValueError, match="CUDA"):
        S._scalar_col(torch.zeros(4), 4, "cpu")


def test_inout_hits_are_checked_before_the_call():
    n = 8
    good = {k: torch.zeros(n, dtype=torch.float64) for k in ("t", "u", "v")}
    good["prim_id"] = torch.zeros(n, dtype=torch.int32)
    good["geom_id"] = torch.zeros(n, dtype=torch.int32)
    bad_len = dict(good, t=torch.zeros(n - 1, dtype=torch.float64))
    with pytest.raises(ValueError, match=r"hits\['t'\]"):
        S._check_hits(bad_len, n, "cpu", torch.float64)
    bad_dtype = dict(good, prim_id=torch.zeros(n, dtype=torch.int64))
    with pytest.raises(ValueError, match=r"hits\['prim_id'\]"):
        S._check_hits(bad_dtype, n, "cpu", torch.float64)
    bad_normal = dict(good, normal=torch.zeros((n, 3), dtype=torch.float64))
    with pytest.raises(ValueError, match=r"hits\['normal'\]"):
        S._check_hits(bad_normal, n, "cpu", torch.float64)
    missing = {k: v for k, v in good.items() if k != "geom_id"}
    with pytest.raises(ValueError, match=r"hits\['geom_id'\]"):
        S._check_hits(missing, n, "cpu", torch.float64)
    # host tensors of the right shape and dtype are still refused: the kernels write through device pointers
    with pytest.raises(ValueError, match="CUDA tensor"):
        S._check_hits(good, n, "cpu", torch.float64)


def test_scene_entry_points_refuse_host_data():
    s = Scene()
    s.add_spheres(np.zeros((1, 3)), [1.0])
    s.commit()
    o = np.zeros((4, 3))
    with pytest.raises(ValueError):
        s.intersect(o, o)
    with pytest.raises(ValueError):
        s.intersect(torch.zeros((4, 3), dtype=torch.float64), torch.zeros((4, 3), dtype=torch.float64))
    with pytest.raises(ValueError):
        s.distance_to_surface(1, torch.zeros((4, 3), dtype=torch.float64), torch.zeros(4, dtype=torch.int32))
    with pytest.raises(ValueError):
        s.closest_surface(o)
    from blazert_amd.mirror import mirror_trace
    with pytest.raises(ValueError):
        mirror_trace(s, o, o)
