"""stride_decode_noalloc, the decode whose kernels take their scratch (the
two-pass walk slots, the work-queue counter) from the library's static
device array and which allocates nothing, against stride_decode bit for
bit: two calls in a row, the second on a smaller stream, in every launch
shape that uses scratch."""
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

GOLD = Path(__file__).resolve().parent / "golden"

SHAPES = [dict(),                                                  # two passes, token chains
          dict(mode=4, tok_chains=0, tok_group=32),                # two passes, token lanes
          dict(mode=2, group_size=16, lds_bytes=24576, grid=64, use_queue=1),   # queue counter
          dict(mode=0, waves=4, stage_words=512, grid=64, use_queue=1)]
SHAPE_IDS = ["chains", "token_lanes", "split_queue", "wave_queue"]


@pytest.mark.parametrize("shape", SHAPES, ids=SHAPE_IDS)
def test_noalloc_decode_equals_decode(shape):
    import torch
    from meshpress_amd import _lib
    from meshpress_amd.utils.paradelta_v5_cuda import ParaDeltaV5GpuDecoder
    L = _lib.lib()
    for name in ("dup_uv_c1", "dup_torus_30x10"):   # the second stream is the smaller one
        data = np.load(GOLD / f"{name}.npz")["data"].tobytes()
        dec = ParaDeltaV5GpuDecoder(data)
        if shape:
            dec.set_tune(**shape)
        v0, t0 = [x.clone() for x in dec.decode()]
        torch.cuda.synchronize()
        V = torch.full_like(v0, float("nan"))
        T = torch.full_like(t0, -1)
        _lib.check(L.stride_decode_noalloc(dec._h, _lib.ptr(V), _lib.ptr(T), _lib.stream_ptr()))
        torch.cuda.synchronize()
        assert torch.equal(V.view(torch.int32), v0.view(torch.int32)), name
        assert torch.equal(T, t0), name


def test_noalloc_decode_rejects_what_it_does_not_cover():
    from meshpress_amd import _lib
    from meshpress_amd.utils.paradelta_v5_cuda import ParaDeltaV5GpuDecoder
    L = _lib.lib()
    dec = ParaDeltaV5GpuDecoder(np.load(GOLD / "dup_one_tri.npz")["data"].tobytes())
    with pytest.raises(ValueError, match="null output"):
        _lib.check(L.stride_decode_noalloc(dec._h, None, None, None))
    v5 = ParaDeltaV5GpuDecoder(np.load(GOLD / "pdlt_one_tri.npz")["data"].tobytes())
    with pytest.raises(ValueError, match="DPR4 streams only"):
        _lib.check(L.stride_decode_noalloc(v5._h, _lib.ptr(v5.d_verts), _lib.ptr(v5.d_tris), None))
