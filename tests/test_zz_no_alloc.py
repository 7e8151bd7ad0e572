"""npde_block_average_noalloc: the block average staged in the library's static device array equals
npde_block_average (which allocates its staging buffer per call) bit for bit, on consecutive calls whose
second is smaller, across instances of a batch; a request larger than the array is refused before any
device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def noalloc(b, inst, m):
    g = b.grids[inst]
    f = g.nx // m
    out = np.full((4, g.ny // f, m), np.nan, dtype=np.float32)
    st = b._lib.npde_block_average_noalloc(b._h, inst, m, out.ctypes.data_as(C.POINTER(C.c_float)))
    return st, out


def test_noalloc_block_average_equals_the_allocating_call(npde_mod):
    from npde.families import riemann
    grids = [npde_mod.make_grid(512, 512, bc_x=1, bc_y=1), npde_mod.make_grid(256, 128, bc_x=1, bc_y=0)]
    with npde_mod.Batch(grids, "hllc", "vanleer", 0.5) as b:
        b.set_state(0, riemann(3, 512)[0])
        b.set_state(1, riemann(4, 256)[0][:, :128, :])
        b.advance_steps(12, 1.0)
        # two calls in a row, the second smaller, then the largest the array holds and another instance
        for inst, m in ((0, 128), (0, 32), (0, 512), (1, 64), (1, 256)):
            st, out = noalloc(b, inst, m)
            assert st == npde_mod._lib.OK, (inst, m, b._lib.npde_last_error(b._h))
            ref = b.block_average(inst, m)
            assert out.view(np.uint32).tobytes() == ref.view(np.uint32).tobytes(), (inst, m)
        # the state moves on: the array keeps nothing from the last call that a new one could read
        b.advance_steps(3, 1.0)
        st, out = noalloc(b, 0, 64)
        assert st == npde_mod._lib.OK and np.array_equal(out, b.block_average(0, 64))


def test_noalloc_block_average_refuses_more_than_the_static_array(npde_mod):
    with npde_mod.Batch([npde_mod.make_grid(1024, 1024, bc_x=1, bc_y=1)], "hllc", "vanleer", 0.5) as b:
        st, out = noalloc(b, 0, 1024)  # 4 x 1024^2 values: four times the array
        assert st == npde_mod._lib.ERR_INVALID
        assert b"exceeds" in b._lib.npde_last_error(b._h)
        assert np.isnan(out).all()  # nothing was written
        st, out = noalloc(b, 0, 512)
        assert st == npde_mod._lib.OK and np.array_equal(out, b.block_average(0, 512))
