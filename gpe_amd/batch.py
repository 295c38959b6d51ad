"""Many independent GP simulations advanced together on one GPU.

`Batch` wraps the gpe_batch_* entry points of include/gpe_hip.h.  Each
instance is a solver() loop of its own -- grid, boundary conditions, order,
trap, dt and number of steps are per instance -- in FP64 or the compensated
FP32 form.  Per stage kind, the instances of one (precision, order, scheme)
run in one persistent launch over a work table of (instance, tile) items;
the step control of every instance (rk4 time bookkeeping, the rk45
accept/reject and dt update of rkqs, solve.f90:180-229) runs on the device
after the final stage, and one iteration is a captured hipGraph.  The
reference has no batch mode (one compile-time grid per job,
parameters.in:23-29); an instance's results equal a `Simulation` run of the
same configuration.
"""
import ctypes
import dataclasses
import json
import os

import numpy as np

from . import _lib
from .config import RunConfig
from .sim import make_params, _ptr

STATUS = {0: "running", 1: "done", -1: "nan", -2: "rejections"}


def footprint(cfg: RunConfig) -> int:
    """Device bytes an instance of this configuration occupies."""
    b = ctypes.c_int64()
    _lib.check(_lib.lib().gpe_batch_footprint(ctypes.byref(make_params(cfg)), ctypes.byref(b)))
    return b.value


class Batch:
    def __init__(self, hbm_budget=0):
        self._L = _lib.lib()
        self._b = ctypes.c_void_p()
        _lib.check(self._L.gpe_batch_create(int(hbm_budget), ctypes.byref(self._b)))
        self.cfgs = {}
        self.targets = {}

    def close(self):
        if self._b:
            self._L.gpe_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- instances ---------------------------------------------------------
    def add(self, cfg: RunConfig, psi, nsteps, dt=None, t=0.0, im_t=0.0):
        """Add an instance; returns its id, or None when it does not fit the
        HBM budget (GPE_EFULL)."""
        if cfg.precision not in ("fp64", "fp32c"):
            raise ValueError("batches run fp64 or fp32c instances")
        psi = np.ascontiguousarray(psi)
        if psi.shape != cfg.shape:
            raise ValueError("psi has shape %s, expected %s" % (psi.shape, cfg.shape))
        if psi.dtype == np.complex64:
            ht = _lib.GPE_HOST_C64
        else:
            psi = psi.astype(np.complex128, copy=False)
            ht = _lib.GPE_HOST_C128
        st = _lib.GpeStepState()
        d = complex(cfg.dt if dt is None else dt)
        st.dt_re, st.dt_im, st.t, st.im_t = d.real, d.imag, t, im_t
        i = ctypes.c_int32()
        rc = self._L.gpe_batch_add(self._b, ctypes.byref(make_params(cfg)), _ptr(psi), ht, ctypes.byref(st),
                                   int(nsteps), ctypes.byref(i))
        if rc == _lib.GPE_EFULL:
            return None
        _lib.check(rc)
        self.cfgs[i.value] = cfg
        self.targets[i.value] = int(nsteps)
        return i.value

    def add_random_phase(self, cfg: RunConfig, seed, nsteps, kmax=0.5, dt=None, t=0.0, im_t=0.0):
        """Add an instance whose initial field is the random-phase state,
        built on the device in the instance's memory; returns its id, or
        None when it does not fit the HBM budget."""
        if cfg.precision not in ("fp64", "fp32c"):
            raise ValueError("batches run fp64 or fp32c instances")
        st = _lib.GpeStepState()
        d = complex(cfg.dt if dt is None else dt)
        st.dt_re, st.dt_im, st.t, st.im_t = d.real, d.imag, t, im_t
        i = ctypes.c_int32()
        rc = self._L.gpe_batch_add_random_phase(self._b, ctypes.byref(make_params(cfg)), ctypes.c_uint64(int(seed)),
                                                float(kmax), ctypes.byref(st), int(nsteps), ctypes.byref(i))
        if rc == _lib.GPE_EFULL:
            return None
        _lib.check(rc)
        self.cfgs[i.value] = cfg
        self.targets[i.value] = int(nsteps)
        return i.value

    def remove(self, iid):
        _lib.check(self._L.gpe_batch_remove(self._b, int(iid)))
        self.cfgs.pop(iid, None)
        self.targets.pop(iid, None)

    # -- running -----------------------------------------------------------
    def run(self, iterations=1):
        """Enqueue iterations (one rk4 step / one rk45 attempt each) for
        every running instance; returns without waiting."""
        _lib.check(self._L.gpe_batch_run(self._b, int(iterations)))

    def sync(self):
        _lib.check(self._L.gpe_batch_sync(self._b))

    def poll(self, cap=4096):
        """[(id, status)] of instances that stopped since the last poll."""
        ids = (ctypes.c_int32 * cap)()
        sts = (ctypes.c_int32 * cap)()
        n = ctypes.c_int32()
        _lib.check(self._L.gpe_batch_poll(self._b, ids, sts, cap, ctypes.byref(n)))
        return [(ids[k], sts[k]) for k in range(n.value)]

    def run_to_completion(self, chunk=16, max_iterations=10 ** 7):
        """Run until every instance has stopped; returns {id: status}."""
        done = {}
        it = 0
        while len(done) < len(self.cfgs) and it < max_iterations:
            self.run(chunk)
            it += chunk
            for i, s in self.poll():
                done[i] = s
        return done

    # -- results -----------------------------------------------------------
    def query(self, iid):
        st = _lib.GpeStepState()
        n = ctypes.c_int64()
        s = ctypes.c_int32()
        nrm = ctypes.c_double()
        pend = ctypes.c_int32()
        _lib.check(self._L.gpe_batch_query(self._b, int(iid), ctypes.byref(st), ctypes.byref(n), ctypes.byref(s),
                                           ctypes.byref(nrm), ctypes.byref(pend)))
        return {"dt": complex(st.dt_re, st.dt_im), "t": st.t, "im_t": st.im_t, "errmax": st.errmax,
                "n_rejects": st.n_rejects, "steps": n.value, "status": s.value, "norm": nrm.value,
                "pending_rejects": pend.value}

    def download(self, iid, dtype=np.complex128):
        cfg = self.cfgs[iid]
        out = np.empty(cfg.shape, dtype=dtype)
        ht = _lib.GPE_HOST_C64 if out.dtype == np.complex64 else _lib.GPE_HOST_C128
        _lib.check(self._L.gpe_batch_download(self._b, int(iid), _ptr(out), ht))
        return out

    def stats(self):
        n, r = ctypes.c_int32(), ctypes.c_int32()
        u, b = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._L.gpe_batch_stats(self._b, ctypes.byref(n), ctypes.byref(r), ctypes.byref(u),
                                           ctypes.byref(b)))
        return {"instances": n.value, "running": r.value, "bytes_in_use": u.value, "budget": b.value}

    # -- launch configuration ------------------------------------------------
    def use_graph(self, enable=True):
        _lib.check(self._L.gpe_batch_use_graph(self._b, 1 if enable else 0))

    def tune_set(self, precision, order, scheme, kind, variant, zchunk):
        _lib.check(self._L.gpe_batch_tune_set(self._b, _lib.PRECISION_CODES[precision], int(order),
                                              _lib.SCHEME_CODES[scheme], _lib.KIND_NAMES.index(kind),
                                              int(variant), int(zchunk)))

    def tune_get(self, precision, order, scheme, kind):
        v, z = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(self._L.gpe_batch_tune_get(self._b, _lib.PRECISION_CODES[precision], int(order),
                                              _lib.SCHEME_CODES[scheme], _lib.KIND_NAMES.index(kind),
                                              ctypes.byref(v), ctypes.byref(z)))
        return v.value, z.value

    def autotune(self, reps=3):
        """Pick tile variant and z-chunk per stage kind by timing the
        batched launches on the current instances; commits no step."""
        _lib.check(self._L.gpe_batch_autotune(self._b, int(reps)))

    # -- checkpoint / restart ------------------------------------------------
    def checkpoint(self, dirpath, ids=None, max_iterations=1000):
        """Write a restart point of the given (default: all running)
        instances: each field in the reference's end_state format
        (io.f90:843-853; complex128, the compensated form as hi + lo) plus a
        manifest with its configuration and step counts.  rk45 instances in
        the middle of a step (after a rejected attempt) are first run to the
        end of that step, so every saved state is a step boundary."""
        from .state_io import write_state
        ids = list(self.cfgs) if ids is None else list(ids)
        for _ in range(max_iterations):
            q = {i: self.query(i) for i in ids}
            if all(v["pending_rejects"] == 0 or v["status"] != 0 for v in q.values()):
                break
            self.run(1)
        else:
            raise RuntimeError("rk45 instances did not reach a step boundary")
        os.makedirs(dirpath, exist_ok=True)
        manifest = {}
        for i in ids:
            v = q[i]
            if v["status"] != 0:
                continue
            fn = "end_state_%d.dat" % i
            write_state(os.path.join(dirpath, fn), self.download(i), p=v["steps"], t=v["t"], im_t=v["im_t"],
                        dt=v["dt"])
            manifest[str(i)] = {"file": fn, "cfg": dataclasses.asdict(self.cfgs[i]), "steps_done": v["steps"],
                                "steps_target": self.targets.get(i, v["steps"])}
        with open(os.path.join(dirpath, "manifest.json"), "w") as fh:
            json.dump(manifest, fh, indent=1)
        return manifest

    def restore(self, dirpath):
        """Re-admit the instances of a checkpoint for their remaining steps;
        returns {old id: new id} (None where the budget is full)."""
        from .state_io import read_state
        with open(os.path.join(dirpath, "manifest.json")) as fh:
            manifest = json.load(fh)
        out = {}
        for old, ent in manifest.items():
            cfg = RunConfig(**ent["cfg"])
            psi, p, t, im_t, dt = read_state(os.path.join(dirpath, ent["file"]))
            out[int(old)] = self.add(cfg, psi, ent["steps_target"] - ent["steps_done"], dt=dt, t=t, im_t=im_t)
        return out
