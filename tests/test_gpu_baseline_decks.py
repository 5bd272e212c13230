"""GPU parity against the CPU oracle on the benchmark decks at full size.

- 64x64 argon CCP (dx 0.5 mm, 100 V, 100 mTorr, 50 particles per cell per
  species, 409 600 particles): three timesteps stage by stage in FP32 and
  FP64, with the checks of test_gpu_stages (bit-exact particles, event
  counts and E field; density within 1e-5; potential within the solver
  tolerance of the oracle's FP64 banded direct solve).
- 256x256 argon CCP (dx 0.2 mm, 200 V, 50 mTorr, 200 per cell per species,
  26.2 M particles): one FP32 timestep stage by stage; its FP64 field
  solution lies within 1e-8 relative of the direct solve.
"""
import pytest

from picmcc import FP32_REFINED, decks
from test_gpu_stages import PRECS, stage_parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,pn", PRECS)
def test_64x64_deck_stage_by_stage(prec, pn):
    stage_parity(decks.config1(), prec, pn, 3)


def test_256x256_deck_one_fp32_step():
    stage_parity(decks.config2(), FP32_REFINED, "f32", 1)
