"""Helpers shared by the GPU tests."""
import numpy as np


def to_oracle(comps):
    """Product components (aq_component fields) -> oracle components."""
    out = []
    for c in comps:
        out.append(dict(kind=c["kind"], nbits=c["num_bits"], is_signed=c["is_signed"],
                        bit_offset=c["bit_offset"], exp_bits=c["exp_bits"],
                        exp_offset=c["exp_offset"], shared=c["shared"],
                        dither=c["use_dithering"], scale=c["scale"], offset=c["offset"]))
    return out


def random_values(comps, n, rng, scale_fixed=0.9):
    """Values inside each component's range (fixed) or spread over magnitudes (float)."""
    vals = np.zeros((len(comps), n), dtype=np.float32)
    for k, c in enumerate(comps):
        if c["kind"] == 0:  # fixed
            if c["is_signed"]:
                lim = c["scale"] * 2 ** (c["num_bits"] - 1)
                vals[k] = c["offset"] + rng.uniform(-scale_fixed, scale_fixed, n) * lim
            else:
                lim = c["scale"] * 2 ** c["num_bits"]
                vals[k] = c["offset"] + rng.uniform(0, scale_fixed, n) * lim
        elif c["kind"] == 1:
            hi = 2 ** (c["num_bits"] - 1) if c["is_signed"] else 2 ** c["num_bits"]
            lo = -hi if c["is_signed"] else 0
            vals[k] = rng.integers(lo, hi, n).astype(np.float32)
        elif c["kind"] == 2:
            mag = 10.0 ** rng.uniform(-6, 3, n)
            sgn = np.where(rng.random(n) < 0.5, -1.0, 1.0) if c["is_signed"] else 1.0
            v = mag * sgn
            v[rng.random(n) < 0.05] = 0.0
            vals[k] = v
        else:
            vals[k] = rng.normal(size=n) * 10
    return vals


def state_layout(kind):
    from autoquantizer_amd import layouts
    from autoquantizer_amd.mpm import plan_state
    if kind == "elastic_2d":
        names, t, g, sh = layouts.elastic_2d()
    elif kind == "hand_2d":
        names, t, g, sh = layouts.hand_scheme_2d()
    else:
        names, t, g, sh = layouts.fluid_3d()
    comps, nw = plan_state(names, t, g, sh)
    return comps, nw
