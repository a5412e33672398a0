"""Clean.clean -- src/imc_clean.jl:6-19: order-preserving removal of killed particles,
executed on the GPU (stream compaction fused with the census tally)."""
from __future__ import annotations

from ._hip import Engine


def clean(particles: Engine) -> None:
    particles.clean()
