"""cwa_grid_build_static: the grid build that allocates nothing (its scratch
is a file-scope device array of the library) against cwa_grid_build, to the
bit, on two calls in a row (the second with fewer particles and cells, so it
reuses the array the first one left behind); sizes beyond the array are
refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _build(cwa, torch, dev, fn, p, pos, stream_sync=True):
    from coupledwateranimation_amd import _lib

    n = pos.shape[0]
    nc = p.grid_dims[0] * p.grid_dims[1] * p.grid_dims[2]
    dpos = torch.tensor(pos, device=dev)
    key = torch.full((max(n, 1),), -1, dtype=torch.int32, device=dev)
    lst = torch.full((max(n, 1),), -1, dtype=torch.int32, device=dev)
    cnt = torch.full((nc,), -1, dtype=torch.int32, device=dev)
    off = torch.full((nc + 1,), -1, dtype=torch.int32, device=dev)
    gmin = (C.c_float * 3)(*p.grid_min)
    dims = (C.c_int32 * 3)(*p.grid_dims)
    _lib.check(getattr(cwa.lib(), fn)(dpos.data_ptr() if n else None, n, gmin, p.cell_size,
                                      dims, key.data_ptr(), cnt.data_ptr(), off.data_ptr(),
                                      lst.data_ptr(), None))
    torch.cuda.synchronize()
    return {"key": key.cpu().numpy()[:n], "count": cnt.cpu().numpy(),
            "offset": off.cpu().numpy(), "list": lst.cpu().numpy()[:n]}


def test_static_grid_build_matches_allocating_build(gpu, cwa):
    import torch
    from coupledwateranimation_amd import configs

    big = configs.dam_break_large(steps=1, ydiv=1)            # 1 048 576 particles
    small = configs.dam_break_small()                          # 4 096 particles
    rng = np.random.default_rng(3)
    pos = np.ones((20_000, 4), np.float32)                     # many in few cells
    pos[:, :3] = rng.uniform(0.1, 0.14, (20_000, 3))
    for name, p, x in (("1M", big.params, big.pos), ("4k", small.params, small.pos),
                       ("clustered", small.params, pos)):
        a = _build(cwa, torch, gpu, "cwa_grid_build_static", p, x)
        b = _build(cwa, torch, gpu, "cwa_grid_build", p, x)
        for k in ("key", "count", "offset", "list"):
            assert np.array_equal(a[k], b[k]), (name, k)


def test_static_grid_build_refuses_oversize(gpu, cwa):
    """More particles or cells than the scratch holds: CWA_EINVAL with the
    reason, and nothing launched (the buffers keep their contents)."""
    import torch
    from coupledwateranimation_amd import configs

    p = configs.dam_break_small().params
    gmin = (C.c_float * 3)(*p.grid_min)
    buf = torch.full((64,), 7, dtype=torch.int32, device=gpu)
    L = cwa.lib()
    ptr = buf.data_ptr()
    rc = L.cwa_grid_build_static(ptr, (1 << 21) + 1, gmin, p.cell_size,
                                 (C.c_int32 * 3)(*p.grid_dims), ptr, ptr, ptr, ptr, None)
    assert rc == -1 and b"static scratch" in L.cwa_last_error()
    rc = L.cwa_grid_build_static(ptr, 0, gmin, p.cell_size, (C.c_int32 * 3)(256, 128, 129),
                                 ptr, ptr, ptr, ptr, None)
    assert rc == -1 and b"static scratch" in L.cwa_last_error()
    torch.cuda.synchronize()
    assert int(buf.min()) == 7 and int(buf.max()) == 7
