"""Statistical parity of the fp32 product path with the oracle (f64).

* Per-pixel means over independent seeds agree within 4 standard errors,
  tested on 4x4-pixel blocks so ~1e-4 Gaussian false alarms are expected
  over all blocks (tolerance: no block beyond 4 SE... one allowed per 2000).
* With common random numbers (same seed, identical per-pixel streams and
  draw order) the relative RMSE between converged 4096-spp images stays
  below 1e-3.
"""
import numpy as np
import pytest

import zigtrace_amd as zt
from oracle import oracle as O

from .conftest import scene_text
from .helpers import block_mean

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,w,h,depth", [("cornell", 64, 64, 16), ("lamps", 64, 36, 12), ("glass", 64, 36, 16)])
def test_seed_means_within_4_standard_errors(name, w, h, depth):
    text = scene_text(name)
    sc = zt.parse(text)
    ref = O.OracleScene.parse(text)
    k, spp = 12, 32
    gpu = np.stack([block_mean(zt.render(sc, zt.Options(w, h, spp, depth, seed=s)), 4) for s in range(k)])
    cpu = np.stack([block_mean(ref.render(w, h, spp, depth, seed=1000 + s)[0], 4) for s in range(k)])
    se = np.sqrt(gpu.var(axis=0, ddof=1) / k + cpu.var(axis=0, ddof=1) / k)
    diff = np.abs(gpu.mean(axis=0) - cpu.mean(axis=0))
    z = np.where(se > 0, diff / np.maximum(se, 1e-300), np.where(diff > 1e-6, np.inf, 0))
    exceed = (z > 4).mean()
    assert exceed <= max(1.0 / 2000, 1.5 / z.size), f"{name}: {exceed:.4f} of block-channels beyond 4 SE"


@pytest.mark.parametrize("name,w,h,depth", [("cornell", 32, 32, 16), ("lamps", 32, 18, 12)])
def test_converged_relative_rmse_with_common_random_numbers(name, w, h, depth):
    text = scene_text(name)
    sc = zt.parse(text)
    ref = O.OracleScene.parse(text)
    got = zt.render(sc, zt.Options(w, h, 4096, depth, seed=11))
    want, _ = ref.render(w, h, 4096, depth, seed=11)
    rmse = np.sqrt(np.mean((got - want) ** 2)) / np.sqrt(np.mean(want ** 2))
    assert rmse < 1e-3, f"{name}: relative RMSE {rmse:.2e}"
