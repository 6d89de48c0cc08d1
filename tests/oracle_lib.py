"""ctypes binding of oracle/liboracle.so -- the CPU checker (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "spatialops-threadfix_amd", "python"))

from nebo_hip import _abi  # noqa: E402

ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")


class GhostT(C.Structure):
    _fields_ = [("minus", _abi.I3), ("plus", _abi.I3)]


_P = C.POINTER
_vp = C.c_void_p
_SIGS = {
    "oracle_node_ghosts": (C.c_int, [_P(_abi.ExprT), C.c_int, C.c_int, _P(GhostT)]),
    "oracle_assign_limits": (C.c_int, [_P(_abi.FieldT), _P(_abi.ExprT), C.c_int, _P(_abi.LimitsT), _P(C.c_int)]),
    "oracle_eval": (C.c_double, [_P(_abi.ExprT), C.c_int, C.c_int, C.c_int, C.c_int]),
    "oracle_assign": (C.c_int, [_P(_abi.FieldT), _P(_abi.ExprT), C.c_int]),
    "oracle_assign_threads": (C.c_int, [_P(_abi.FieldT), _P(_abi.ExprT), C.c_int, C.c_int]),
    "oracle_masked_assign": (C.c_int, [_P(_abi.FieldT), _vp, C.c_int, _P(_abi.ExprT), C.c_int]),
    "oracle_reduce": (C.c_int, [_P(_abi.ExprT), C.c_int, C.c_int, _P(C.c_double)]),
    "oracle_bc_apply": (C.c_int, [_P(_abi.FieldT), _P(_abi.ExprT), _P(_abi.StencilT), _vp, _vp, C.c_int,
                                  _vp, C.c_int, C.c_int]),
    "oracle_fill_ghosts_periodic": (None, [_P(_abi.FieldT)]),
    "oracle_mask_bits": (None, [_P(_abi.MaskT), _vp, C.c_int, _vp]),
    "oracle_divkgrad": (None, [_P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.AxisCoefsT), C.c_int]),
    "oracle_divkgrad_threads": (None, [_P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.AxisCoefsT),
                                       C.c_int, C.c_int]),
    "oracle_lapl_euler": (None, [_P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.AxisCoefsT), C.c_double]),
    "oracle_diffusion_step": (None, [_P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.AxisCoefsT),
                                     C.c_double]),
    "oracle_diffusion_adjoint_step": (None, [_P(_abi.FieldT), _vp, _P(_abi.FieldT), _P(_abi.FieldT),
                                             _P(_abi.FieldT), _P(_abi.AxisCoefsT), C.c_double]),
    "oracle_transport_stage": (None, [_P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.FieldT),
                                      _P(_abi.FieldT), _P(_abi.AxisCoefsT), C.c_double, C.c_int]),
    "oracle_transport_stage_threads": (None, [_P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.FieldT),
                                              _P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.AxisCoefsT), C.c_double,
                                              C.c_int, C.c_int]),
    "oracle_transport_adjoint_stage": (None, [_P(_abi.FieldT), _vp, _vp, _P(_abi.FieldT), _vp, _P(_abi.FieldT),
                                              _P(_abi.FieldT), _P(_abi.FieldT), _P(_abi.AxisCoefsT), C.c_double,
                                              C.c_int]),
    "oracle_particle_to_cell": (C.c_int, [_P(_abi.FieldT), _vp, _vp, _vp, _vp, _vp, C.c_int64,
                                          _P(_abi.ParticleGeomT)]),
    "oracle_cell_to_particle": (C.c_int, [_vp, _P(_abi.FieldT), _vp, _vp, _vp, _vp, C.c_int64,
                                          _P(_abi.ParticleGeomT)]),
    "oracle_point_to_field": (None, [_vp, _vp, _vp, C.c_int64]),
    "oracle_field_to_point": (None, [_vp, _vp, _vp, C.c_int64]),
    "oracle_splitmix64": (C.c_uint64, [_P(C.c_uint64)]),
    "oracle_inv_erf": (C.c_double, [C.c_double]),
    "oracle_inv_erfc": (C.c_double, [C.c_double]),
    "oracle_fill_uniform": (None, [_vp, C.c_size_t, C.c_uint64, C.c_double, C.c_double]),
    "oracle_fill_normal": (None, [_vp, C.c_size_t, C.c_uint64]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"oracle not built: {ORACLE_PATH} (run __graft_entry__.build())")
        L = C.CDLL(ORACLE_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle status {rc}")


def assign(lhs, rhs_expr, use_ghost=True, threads=0):
    """lhs <<= rhs on host copies (lhs is a nebo_hip.Field with a host array)."""
    from nebo_hip.expr import serialize, E
    s = serialize(E(rhs_expr), "host")
    d = lhs.desc("host")
    if threads:
        check(lib().oracle_assign_threads(C.byref(d), s.ref, int(use_ghost), int(threads)))
    else:
        check(lib().oracle_assign(C.byref(d), s.ref, int(use_ghost)))
    return lhs


def limits(lhs, rhs_expr, use_ghost=True):
    from nebo_hip.expr import serialize, E
    s = serialize(E(rhs_expr), "host")
    d = lhs.desc("host")
    lim = _abi.LimitsT()
    empty = C.c_int(0)
    check(lib().oracle_assign_limits(C.byref(d), s.ref, int(use_ghost), C.byref(lim), C.byref(empty)))
    return None if empty.value else (tuple(lim.lo), tuple(lim.hi))


def reduce(expr, op, use_ghost=True):
    from nebo_hip.expr import serialize, E
    s = serialize(E(expr), "host")
    out = C.c_double(0)
    check(lib().oracle_reduce(s.ref, int(op), int(use_ghost), C.byref(out)))
    return out.value


def masked_assign(mask, lhs, rhs_expr, passes=2):
    from nebo_hip.expr import serialize, E
    import numpy as np
    s = serialize(E(rhs_expr), "host")
    d = lhs.desc("host")
    pts = np.ascontiguousarray(mask.points.reshape(-1), dtype=np.int32)
    check(lib().oracle_masked_assign(C.byref(d), pts.ctypes.data, len(mask.points), s.ref, passes))
    return lhs


def masked_assign_points(points, lhs, rhs_expr, passes=1):
    from nebo_hip.expr import serialize, E
    s = serialize(E(rhs_expr), "host")
    d = lhs.desc("host")
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.int32).reshape(-1))
    check(lib().oracle_masked_assign(C.byref(d), pts.ctypes.data, len(pts) // 3, s.ref, passes))
    return lhs


def bc_apply(op, mask_points, phi, gamma, minus):
    """NeboBoundaryConditionBuilder<op>(mask, phi, gamma, minus) on host copies.
    ``op`` is a nebo_hip Stencil with src = phi location, dst = gamma location."""
    from nebo_hip.expr import serialize, E
    from nebo_hip.fields import LOCATIONS
    s = serialize(E(gamma), "host")
    d = phi.desc("host")
    st = op.struct()
    po = (C.c_int32 * 3)(*LOCATIONS[op.src][0])
    go = (C.c_int32 * 3)(*LOCATIONS[op.dst][0])
    fd = -1 if op.src != op.dst else int(op.axis)
    pts = np.ascontiguousarray(np.asarray(mask_points, dtype=np.int32).reshape(-1))
    check(lib().oracle_bc_apply(C.byref(d), s.ref, C.byref(st), C.addressof(po), C.addressof(go), fd,
                                pts.ctypes.data, len(pts) // 3, int(bool(minus))))
    return phi


def _arr(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def particle_to_cell(dest, src, px, py, pz, psize, geom):
    """ParticleToCell::apply_to_field on host copies; geom = (dx, xlo, dy, ylo, dz, zlo)."""
    keep = [_arr(v) for v in (src, px, py, pz, psize)]
    ptr = [None if v is None else v.ctypes.data for v in keep]
    g = _abi.ParticleGeomT(*[float(v) for v in geom])
    d = dest.desc("host")
    check(lib().oracle_particle_to_cell(C.byref(d), ptr[0], ptr[1], ptr[2], ptr[3], ptr[4], len(keep[0]),
                                        C.byref(g)))
    return dest


def cell_to_particle(src, px, py, pz, psize, geom, n):
    keep = [_arr(v) for v in (px, py, pz, psize)]
    ptr = [None if v is None else v.ctypes.data for v in keep]
    out = np.zeros(n)
    g = _abi.ParticleGeomT(*[float(v) for v in geom])
    d = src.desc("host")
    check(lib().oracle_cell_to_particle(out.ctypes.data, C.byref(d), ptr[0], ptr[1], ptr[2], ptr[3], n,
                                        C.byref(g)))
    return out


def point_to_field(values, src, idx):
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    src = _arr(src)
    lib().oracle_point_to_field(values.ctypes.data, src.ctypes.data, idx.ctypes.data, len(idx))


def field_to_point(values, idx):
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = np.zeros(len(idx))
    lib().oracle_field_to_point(out.ctypes.data, values.ctypes.data, idx.ctypes.data, len(idx))
    return out


def fill_periodic(f):
    d = f.desc("host")
    lib().oracle_fill_ghosts_periodic(C.byref(d))
    return f


def divkgrad(r, k, phi, coefs, form=0, threads=0):
    dr, dk, dp = r.desc("host"), k.desc("host"), phi.desc("host")
    if threads:
        lib().oracle_divkgrad_threads(C.byref(dr), C.byref(dk), C.byref(dp), C.byref(coefs), form, threads)
    else:
        lib().oracle_divkgrad(C.byref(dr), C.byref(dk), C.byref(dp), C.byref(coefs), form)
    return r


def lapl_euler(out, phi, coefs, s):
    do, dp = out.desc("host"), phi.desc("host")
    lib().oracle_lapl_euler(C.byref(do), C.byref(dp), C.byref(coefs), float(s))
    return out


def diffusion_step(out, phi, k, coefs, dt):
    do, dp, dk = out.desc("host"), phi.desc("host"), k.desc("host")
    lib().oracle_diffusion_step(C.byref(do), C.byref(dp), C.byref(dk), C.byref(coefs), float(dt))
    return out


def diffusion_adjoint_step(lam_out, gk, lam, phi, k, coefs, dt):
    dlo, dl, dp, dk = lam_out.desc("host"), lam.desc("host"), phi.desc("host"), k.desc("host")
    lib().oracle_diffusion_adjoint_step(C.byref(dlo), gk.ctypes.data, C.byref(dl), C.byref(dp), C.byref(dk),
                                        C.byref(coefs), float(dt))
    return lam_out


def _face_array(us):
    arr = (_abi.FieldT * 3)()
    for d in range(3):
        arr[d] = us[d].desc("host")
    return arr


def transport_stage(out, phin, s, k, us, coefs, dt, stage, threads=0):
    """One SSP-RK3 stage of the transport right-hand side (fills out's periodic ghosts);
    threads > 0 splits the interior into z slabs (same bits)."""
    do, dn, ds, dk = out.desc("host"), phin.desc("host"), s.desc("host"), k.desc("host")
    if threads:
        lib().oracle_transport_stage_threads(C.byref(do), C.byref(dn), C.byref(ds), C.byref(dk),
                                             _face_array(us), C.byref(coefs), float(dt), int(stage), int(threads))
    else:
        lib().oracle_transport_stage(C.byref(do), C.byref(dn), C.byref(ds), C.byref(dk), _face_array(us),
                                     C.byref(coefs), float(dt), int(stage))
    return out


def transport_adjoint_stage(out, acc_out, gk, a, acc_in, s, k, us, coefs, dt, stage):
    """Transpose of one stage; acc_out / acc_in / gk are numpy arrays of the field's size
    (acc_out unused for stage 1, acc_in unused for stage 3)."""
    do, da, ds, dk = out.desc("host"), a.desc("host"), s.desc("host"), k.desc("host")
    ptr = lambda x: None if x is None else x.ctypes.data
    lib().oracle_transport_adjoint_stage(C.byref(do), ptr(acc_out), ptr(gk), C.byref(da), ptr(acc_in),
                                         C.byref(ds), C.byref(dk), _face_array(us), C.byref(coefs),
                                         float(dt), int(stage))
    return out


def transport_forward_cpu(L, phi0, k, us, coefs, dt, steps):
    """SSP-RK3 trajectory through the oracle's stages.  Returns the per-step
    (phi_n, s1, s2) states (host arrays) and phi_T."""
    from nebo_hip import Field
    hf = lambda a: Field(L, host=np.asarray(a, dtype=np.float64).copy())
    states, phi = [], np.asarray(phi0, dtype=np.float64).copy()
    for _ in range(steps):
        s1, s2, nxt = hf(np.zeros(L.npts)), hf(np.zeros(L.npts)), hf(np.zeros(L.npts))
        transport_stage(s1, hf(phi), hf(phi), k, us, coefs, dt, 1)
        transport_stage(s2, hf(phi), s1, k, us, coefs, dt, 2)
        transport_stage(nxt, hf(phi), s2, k, us, coefs, dt, 3)
        states.append((phi, s1.host, s2.host))
        phi = nxt.host
    return states, phi


def transport_backward_cpu(L, lamT, k, us, coefs, dt, states):
    """Reverse sweep through the oracle's adjoint stages: (dL/dphi_0, dL/dk)."""
    from nebo_hip import Field
    hf = lambda a: Field(L, host=np.asarray(a, dtype=np.float64).copy())
    gk = np.zeros(L.npts)
    lam = np.asarray(lamT, dtype=np.float64).copy()
    for phin, s1, s2 in reversed(states):
        acc = np.zeros(L.npts)
        mu2, mu1, out = hf(np.zeros(L.npts)), hf(np.zeros(L.npts)), hf(np.zeros(L.npts))
        transport_adjoint_stage(mu2, acc, gk, hf(lam), None, hf(s2), k, us, coefs, dt, 3)
        acc2 = np.zeros(L.npts)
        transport_adjoint_stage(mu1, acc2, gk, mu2, acc, hf(s1), k, us, coefs, dt, 2)
        transport_adjoint_stage(out, None, gk, mu1, acc2, hf(phin), k, us, coefs, dt, 1)
        lam = out.host
    return lam, gk


def ulp_distance(a, b):
    """Representable doubles between a and b, elementwise (the count
    boost::math::float_distance gives, which the reference's ulp comparisons in
    spatialops/structured/FieldComparisons.h use): 0 for equal values (and for two
    NaNs), inf when exactly one is NaN."""
    a = np.atleast_1d(np.asarray(a, dtype=np.float64))
    b = np.atleast_1d(np.asarray(b, dtype=np.float64))
    ia, ib = a.view(np.int64), b.view(np.int64)
    lowest = np.int64(-2 ** 63)
    with np.errstate(over="ignore"):
        oa = np.where(ia < 0, lowest - ia, ia)
        ob = np.where(ib < 0, lowest - ib, ib)
        same_side = (oa >= 0) == (ob >= 0)
        d = np.where(same_side, np.abs(oa - ob).astype(np.float64),
                     np.abs(oa.astype(np.float64) - ob.astype(np.float64)))
    na, nb = np.isnan(a), np.isnan(b)
    return np.where(na & nb, 0.0, np.where(na | nb, np.inf, d))


def golden_special(name):
    """(inputs, correctly rounded outputs) of tests/golden/special_functions.json."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "special_functions.json")
    pairs = json.load(open(path))[name]
    return (np.array([float.fromhex(x) for x, _ in pairs]), np.array([float.fromhex(v) for _, v in pairs]))
