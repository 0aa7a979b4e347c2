"""ctypes bindings of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module.  It wraps ``liboracle.so`` built from
``deb_oracle.cpp`` (a plain C++ restatement of the reference's
``evolve_perturbations`` path; see that file's header for the reference
file:line map).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(force: bool = False) -> str:
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True, capture_output=True)
    return path


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def _p(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


class Problem:
    """One cosmology's packed inputs for the oracle (reals only)."""

    def __init__(self, param, cfg, b: int = 0):
        import sys
        pkg = os.path.join(os.path.dirname(_HERE), "disco-eb_amd")
        if pkg not in sys.path:
            sys.path.insert(0, pkg)
        from discoeb_amd import packing
        self.cfg = cfg
        self.cosmo = np.ascontiguousarray(packing.pack_cosmo(param)[b])
        tab, desc = packing.pack_tables(param)
        self.tables = np.ascontiguousarray(tab[b])
        self.desc = np.ascontiguousarray(desc)
        self.q, self.w = packing.nu_nodes(cfg.nqmax)
        if self.q.size == 0:
            self.q = np.zeros(1)
            self.w = np.zeros(1)
        self.n = packing.nvar(cfg.lmaxg, cfg.lmaxgp, cfg.lmaxr, cfg.lmaxnu, cfg.nqmax)

    def _common(self):
        return (ctypes.byref(self.cfg), _p(self.cosmo), _p(self.tables), _p(self.desc, _ip), _p(self.q), _p(self.w))

    def rhs(self, k, tau, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        f = np.zeros(self.n)
        lib().oracle_rhs(*self._common(), ctypes.c_double(k), ctypes.c_double(tau), _p(y), _p(f))
        return f

    def jacobian(self, k, tau, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        J = np.zeros((self.n, self.n))
        dT = np.zeros(self.n)
        lib().oracle_jacobian(*self._common(), ctypes.c_double(k), ctypes.c_double(tau), _p(y), _p(J), _p(dT))
        return J, dT

    def rodas5_step(self, k, t0, t1, y0):
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        y1 = np.zeros(self.n)
        err = np.zeros(self.n)
        lib().oracle_rodas5_step(*self._common(), ctypes.c_double(k), ctypes.c_double(t0), ctypes.c_double(t1),
                                 _p(y0), _p(y1), _p(err))
        return y1, err

    def ics(self, k, tau_out_min):
        ts = ctypes.c_double(0.0)
        y0 = np.zeros(self.n)
        lib().oracle_ics(*self._common(), ctypes.c_double(k), ctypes.c_double(tau_out_min), ctypes.byref(ts), _p(y0))
        return ts.value, y0

    def ics_at(self, k, tau):
        y0 = np.zeros(self.n)
        lib().oracle_ics_at(*self._common(), ctypes.c_double(k), ctypes.c_double(tau), _p(y0))
        return y0

    def outputs(self, k, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        o = np.zeros(20)
        lib().oracle_outputs(*self._common(), ctypes.c_double(k), _p(y), _p(o))
        return o

    def evolve(self, kmodes, tau_out, nthreads: int = 1):
        kmodes = np.ascontiguousarray(kmodes, dtype=np.float64)
        tau_out = np.ascontiguousarray(np.atleast_1d(tau_out), dtype=np.float64)
        nk, nout = kmodes.size, tau_out.size
        width = self.n if self.cfg.return_full else 20
        y = np.zeros((nk, nout, width))
        st = np.zeros(nk, dtype=np.int32)
        ns = np.zeros(nk, dtype=np.int32)
        na = np.zeros(nk, dtype=np.int32)
        ts = np.zeros(nk)
        sec = np.zeros(nk)
        lib().oracle_evolve(*self._common(), ctypes.c_int(nk), _p(kmodes), ctypes.c_int(nout), _p(tau_out), _p(y),
                            _p(st, _ip), _p(ns, _ip), _p(na, _ip), _p(ts), ctypes.c_int(nthreads), _p(sec))
        return dict(y=y, status=st, nsteps=ns, naccept=na, tau_start=ts, seconds=sec)

    def evolve_batched(self, kmodes, tau_out, batch: int, nthreads: int = 1):
        """evolve_modes_batched: groups of `batch` consecutive k share one step sequence."""
        kmodes = np.ascontiguousarray(kmodes, dtype=np.float64)
        tau_out = np.ascontiguousarray(np.atleast_1d(tau_out), dtype=np.float64)
        nk, nout = kmodes.size, tau_out.size
        width = self.n if self.cfg.return_full else 20
        y = np.zeros((nk, nout, width))
        st = np.zeros(nk, dtype=np.int32)
        ns = np.zeros(nk, dtype=np.int32)
        na = np.zeros(nk, dtype=np.int32)
        r = lib().oracle_evolve_batched(*self._common(), ctypes.c_int(nk), _p(kmodes), ctypes.c_int(nout),
                                        _p(tau_out), ctypes.c_int(batch), _p(y), _p(st, _ip), _p(ns, _ip),
                                        _p(na, _ip), ctypes.c_int(nthreads))
        if r != 0:
            raise ValueError("num_k must be a multiple of batch")
        return dict(y=y, status=st, nsteps=ns, naccept=na)

    def step_control(self, k, err, y0, y1, prev_inv=1.0, prev_prev_inv=1.0):
        """(norm, accepted, factor, inverse norm carried on) of one attempted step."""
        err = np.ascontiguousarray(err, dtype=np.float64)
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        y1 = np.ascontiguousarray(y1, dtype=np.float64)
        out = np.zeros(4)
        lib().oracle_step_control(ctypes.byref(self.cfg), ctypes.c_double(k), _p(err), _p(y0), _p(y1),
                                  ctypes.c_double(prev_inv), ctypes.c_double(prev_prev_inv), _p(out))
        return out[0], bool(out[1]), out[2], out[3]

    def batch_trace(self, kmodes, tau_out, max_trace: int = 8):
        """The modes ``kmodes`` as one group of evolve_modes_batched: (attempted
        steps, trace[i] = (t0, t1, group error norm) of the first steps)."""
        kmodes = np.ascontiguousarray(kmodes, dtype=np.float64)
        tau_out = np.ascontiguousarray(np.atleast_1d(tau_out), dtype=np.float64)
        tr = np.zeros((max_trace, 3))
        ns = lib().oracle_batch_trace(*self._common(), ctypes.c_int(kmodes.size), _p(kmodes),
                                      ctypes.c_int(tau_out.size), _p(tau_out), ctypes.c_int(max_trace), _p(tr))
        return ns, tr

    # ---- tangents along one input direction (dcos[24], dtab[stride]) ----
    def _tan(self, dcos, dtab):
        dcos = np.ascontiguousarray(dcos, dtype=np.float64)
        dtab = np.ascontiguousarray(dtab, dtype=np.float64)
        self._keep = (dcos, dtab)
        return (ctypes.byref(self.cfg), _p(self.cosmo), _p(self.tables), _p(dcos), _p(dtab), _p(self.desc, _ip),
                _p(self.q), _p(self.w))

    def rhs_tangent(self, k, tau, dtau, y, dy, dcos, dtab):
        y = np.ascontiguousarray(y, dtype=np.float64)
        dy = np.ascontiguousarray(dy, dtype=np.float64)
        f, df = np.zeros(self.n), np.zeros(self.n)
        lib().oracle_rhs_tangent(*self._tan(dcos, dtab), ctypes.c_double(k), ctypes.c_double(tau),
                                 ctypes.c_double(dtau), _p(y), _p(dy), _p(f), _p(df))
        return f, df

    def step_tangent(self, k, t0, dt0, t1, dt1, y0, dy0, dcos, dtab):
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        dy0 = np.ascontiguousarray(dy0, dtype=np.float64)
        y1, dy1 = np.zeros(self.n), np.zeros(self.n)
        lib().oracle_rodas5_step_tangent(*self._tan(dcos, dtab), ctypes.c_double(k), ctypes.c_double(t0),
                                         ctypes.c_double(dt0), ctypes.c_double(t1), ctypes.c_double(dt1), _p(y0),
                                         _p(dy0), _p(y1), _p(dy1))
        return y1, dy1

    def outputs_tangent(self, k, y, dy, dcos, dtab):
        y = np.ascontiguousarray(y, dtype=np.float64)
        dy = np.ascontiguousarray(dy, dtype=np.float64)
        o, do = np.zeros(20), np.zeros(20)
        lib().oracle_outputs_tangent(*self._tan(dcos, dtab), ctypes.c_double(k), _p(y), _p(dy), _p(o), _p(do))
        return o, do

    def ics_tangent(self, k, tmin, dtmin, dcos, dtab):
        ts, dts = ctypes.c_double(0.0), ctypes.c_double(0.0)
        y0, dy0 = np.zeros(self.n), np.zeros(self.n)
        lib().oracle_ics_tangent(*self._tan(dcos, dtab), ctypes.c_double(k), ctypes.c_double(tmin),
                                 ctypes.c_double(dtmin), ctypes.byref(ts), ctypes.byref(dts), _p(y0), _p(dy0))
        return ts.value, dts.value, y0, dy0

    def evolve_tangent(self, kmodes, tau_out, dtau_out, dcos, dtab, nthreads: int = 1):
        """20 output fields and their tangent along (dcos, dtab, dtau_out), with the
        accepted steps of the plain solve held fixed as fractions of [t0, t1]."""
        kmodes = np.ascontiguousarray(kmodes, dtype=np.float64)
        tau_out = np.ascontiguousarray(np.atleast_1d(tau_out), dtype=np.float64)
        dtau_out = np.ascontiguousarray(np.atleast_1d(dtau_out), dtype=np.float64)
        nk, nout = kmodes.size, tau_out.size
        y, dy = np.zeros((nk, nout, 20)), np.zeros((nk, nout, 20))
        st = np.zeros(nk, dtype=np.int32)
        na = np.zeros(nk, dtype=np.int32)
        sec = np.zeros(nk)
        lib().oracle_evolve_tangent(*self._tan(dcos, dtab), ctypes.c_int(nk), _p(kmodes), ctypes.c_int(nout),
                                    _p(tau_out), _p(dtau_out), _p(y), _p(dy), _p(st, _ip), _p(na, _ip),
                                    ctypes.c_int(nthreads), _p(sec))
        return dict(y=y, dy=dy, status=st, naccept=na, seconds=sec)

    def replay_tangent(self, k, times, tau_out, dtau_out, dcos, dtab):
        """Output fields and their tangent for the accepted step end times
        ``times`` (times[0] = start time) of some other solve of this mode."""
        times = np.ascontiguousarray(times, dtype=np.float64)
        tau_out = np.ascontiguousarray(np.atleast_1d(tau_out), dtype=np.float64)
        dtau_out = np.ascontiguousarray(np.atleast_1d(dtau_out), dtype=np.float64)
        nout = tau_out.size
        y, dy = np.zeros((nout, 20)), np.zeros((nout, 20))
        lib().oracle_replay_tangent(*self._tan(dcos, dtab), ctypes.c_double(k), ctypes.c_int(times.size), _p(times),
                                    ctypes.c_int(nout), _p(tau_out), _p(dtau_out), _p(y), _p(dy))
        return y, dy

    def perturbed(self, eps, dcos, dtab):
        """A copy whose inputs are moved by eps along (dcos, dtab)."""
        other = Problem.__new__(Problem)
        other.__dict__.update(self.__dict__)
        other.cosmo = np.ascontiguousarray(self.cosmo + eps * np.asarray(dcos))
        other.tables = np.ascontiguousarray(self.tables + eps * np.asarray(dtab))
        return other
