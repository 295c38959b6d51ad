"""Persisted launch configurations of the stage kernels.

The autotuner (gpe_autotune in the C ABI) times every compiled tile variant
and several z-chunk lengths per stage kind on the resident field.  Results
are kept in a JSON file keyed by (grid, order, precision, scheme, boundary
condition, fusion) so a later run of the same problem skips the search; the
library's gpe_tune_load applies an entry to a context (also from C and from
the Fortran drop-in).
"""
import json
import os

DEFAULT_PATH = os.environ.get("GPE_TUNING_FILE", os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                              "tuning.json"))


def key(cfg, fused=None):
    """The table key of cfg's problem, as gpe_tune_key forms it for a
    context.  fused: whether the context allows fused stage pairs (default:
    the library's default, allowed for the FP32 forms)."""
    if fused is None:
        fused = cfg.precision != "fp64"
    return "%dx%dx%d/o%d/%s/%s/b%d%s" % (cfg.nx, cfg.ny, cfg.nz, cfg.order, cfg.precision, cfg.scheme, cfg.bcs,
                                         "/fused" if fused else "")


def load(path=DEFAULT_PATH):
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def save(table, path=DEFAULT_PATH):
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(table, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(sim, path=DEFAULT_PATH):
    """Apply a stored configuration for sim's problem; True if one existed.
    The library reads the table (gpe_tune_load), so C and Fortran callers
    apply the same entries."""
    return sim.tune_load(path)


def tune(sim, reps=3, path=DEFAULT_PATH, store=True):
    """Autotune sim's stage kernels (state unchanged) and persist the result."""
    got = sim.autotune(reps)
    ent = {k: [v[0], v[1]] for k, v in got.items()}
    if store:
        table = load(path)
        table[sim.tune_key()] = ent
        save(table, path)
    return ent
