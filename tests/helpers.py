"""Shared helpers of the parity tests: drive the product (through the C ABI) and
the oracle from the same descriptor and copy state between them."""
import numpy as np

from ariel_amd import _abi

FIELDS = [_abi.FIELD_UX, _abi.FIELD_UY, _abi.FIELD_UZ, _abi.FIELD_DUX, _abi.FIELD_DUY,
          _abi.FIELD_DUZ, _abi.FIELD_P, _abi.FIELD_D, _abi.FIELD_L]


def copy_state(src_oracle, dst_gpu):
    """Load the oracle's particles, grids, flags and max density into the GPU sim,
    then sort there (the GPU keeps particles sorted; the next stage needs that)."""
    pos, vel, rho = src_oracle.get_particles()
    dst_gpu.set_particles(pos, vel, rho)
    for f in FIELDS:
        dst_gpu.set_field(f, src_oracle.get_field(f))
    dst_gpu.set_flags(src_oracle.get_flags())
    dst_gpu.max_density = src_oracle.max_density
    dst_gpu.run_stage(_abi.STAGE_SORT)


def assert_bitexact(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))) if a.dtype.kind == "f"
                         else a != b)
    assert bad.size == 0, (f"{what}: {bad.size} of {a.size} differ; first at {bad[:5]}: "
                           f"{a.ravel()[bad[:5]]} vs {b.ravel()[bad[:5]]}")


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return float(np.abs(a - b).max() / scale)
