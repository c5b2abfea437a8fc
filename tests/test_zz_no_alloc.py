"""dg_gpu_mass_noalloc (the entry point that allocates nothing) against dg_gpu_mass, through the
C ABI: bit for bit on two calls in a row, the second on a smaller mesh, so the second call reads
partial sums that the first left in the shared working array and must not be fooled by them."""
import ctypes as C

import pytest

from dgswem_amd import gpu, mesh

pytestmark = pytest.mark.gpu


def _call(fn, h):
    m = C.c_double(float("nan"))
    rc = fn(h, C.byref(m))
    if rc != 0:
        raise gpu.GpuError(gpu.load_library().dg_gpu_last_error().decode())
    return m.value


@pytest.mark.parametrize("p", [1, 2])
def test_noalloc_mass_equals_mass_on_two_calls(p):
    L = gpu.load_library()
    # 48x24 squares (2304 elements, 10 blocks) then 10x6 (120 elements, one partial block)
    for nx, ny, steps in ((48, 24, 30), (10, 6, 7)):
        case = mesh.structured_basin(nx=nx, ny=ny, p=p, slopeflag=5, dt=2.0 if p == 2 else 5.0,
                                     nsteps=steps)
        g = gpu.GpuSolver(case)
        g.step(steps)
        print(f"no-alloc mass: p={p} {2 * nx * ny} elements after {steps} steps", flush=True)
        want = _call(L.dg_gpu_mass, g.h)
        got = _call(L.dg_gpu_mass_noalloc, g.h)
        assert got == want, (got, want)
        # the second call on the same handle reads what the first left behind and agrees again
        assert _call(L.dg_gpu_mass_noalloc, g.h) == want
        g.close()
