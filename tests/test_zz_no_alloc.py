"""oc_energy_noalloc (no device allocation: the reduction records live in the library's fixed device array)
against oc_energy through the C ABI: the same kernels and reduction order, so the results are bit-identical, on two
calls in a row with the second on a smaller mesh (the array keeps the first call's records), both weightings."""
import numpy as np
import pytest

from optcuts_amd import capi
from optcuts_amd.meshgen import height_field_disc

from _helpers import distort

pytestmark = pytest.mark.gpu


def test_energy_noalloc_matches_energy_on_two_calls():
    ctx = capi.Context(0)
    try:
        for m in (181, 23):  # 64,800 then 968 triangles
            V, F, UV = height_field_disc(m, seed=m)
            ctx.set_mesh(V, F, fixed=(0,))
            ctx.set_uv(distort(UV, m, 0.05))
            for uni in (False, True):
                ref = ctx.energy(uni)
                got = [ctx.energy_noalloc(uni) for _ in range(2)]
                print(m, uni, ref, got, flush=True)
                assert np.isfinite(ref) and ref > 0
                assert got[0] == ref and got[1] == ref
    finally:
        ctx.close()
