"""The batch step that allocates nothing (fx_hydro_advance_batch_noalloc)
against fx_hydro_advance_batch, through the C ABI: the same bits, on two
calls in a row, the second with fewer blocks (the fixed scratch keeps the
first call's contents, which must not leak into the second), and the
fixed scratch's size limit refused from the host.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from flash5_amd import workloads as W
from flash5_amd.hydro import FxError, Hydro, HydroParams

pytestmark = pytest.mark.gpu


def _blocks(hp: HydroParams, nblocks: int, seed: int) -> torch.Tensor:
    U = W.fourier_blocks(nblocks, n=hp.nxb, gamma=hp.gamma, seed=seed)
    if hp.ndim == 2:
        k0 = hp.ng + hp.nxb // 3
        U = U[:, :, k0:k0 + 1].contiguous()
    return U


def _both(hp: HydroParams, U0: torch.Tensor, dx: float):
    deltas = (dx, dx, dx)
    dt = W.cfl_dt(U0, hp.ndim, hp.ng, deltas, 0.5)
    hy = Hydro(hp)
    ref = U0.to("cuda:0")
    st_ref = hy.advance_batch(ref, dt, deltas)
    out = U0.to("cuda:0")
    st = hy.advance_batch_noalloc(out, dt, deltas)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy(), st_ref, st


@pytest.mark.parametrize("ndim,n,riemann,sizes", [
    (3, 16, "hllc", (77, 5)),      # 77 blocks of 16^3 fill the fixed scratch
    (2, 8, "hll", (300, 7)),
])
def test_noalloc_matches_advance_batch_on_two_calls(ndim, n, riemann, sizes):
    hp = HydroParams(ndim=ndim, nxb=n, nyb=n, nzb=n if ndim == 3 else 1, gamma=5.0 / 3.0,
                     RiemannSolver=riemann)
    for call, nb in enumerate(sizes):
        print(f"no-alloc call {call}: ndim={ndim} n={n} {riemann} blocks={nb}", flush=True)
        U0 = _blocks(hp, nb, seed=40 + call)
        ref, out, st_ref, st = _both(hp, U0, 1.0 / n)
        assert st.code == 0 and st_ref.code == 0
        assert st.fallbacks == st_ref.fallbacks
        assert np.array_equal(out, ref), f"call {call}: differs from fx_hydro_advance_batch"
        assert not np.array_equal(out, U0.numpy())


def test_noalloc_refuses_more_than_its_scratch():
    hp = HydroParams(ndim=3, nxb=16, nyb=16, nzb=16, gamma=5.0 / 3.0)
    hy = Hydro(hp)
    U = torch.zeros((78,) + tuple(_blocks(hp, 1, 0).shape[1:]), dtype=torch.float64,
                    device="cuda:0")
    with pytest.raises(FxError) as e:
        hy.advance_batch_noalloc(U, 0.01, (1 / 16,) * 3)
    assert e.value.status.code == 3
    assert float(U.abs().max()) == 0.0      # refused before any kernel ran
