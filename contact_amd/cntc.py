"""ctypes mirror of the reference's python_intfc over libcntc_amd.so.

Function names, arguments and return values follow python_intfc/cntc_*.py (module 3);
arrays come back as numpy arrays of shape (my, mx) like the reference wrapper.  The library
must have been built (contact_amd/build_lib.py); loading fails loudly otherwise.
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libcntc_amd.so")

CNTC = dict(if_units=1933, un_cntc=1934, un_spck=1935, un_si=1936, ic_config=1967,
            ic_pvtime=1970, ic_bound=1971, ic_tang=1972, ic_norm=1973, ic_force=1974,
            ic_frclaw=1976, ic_discns=1977, ic_inflcf=1978, ic_mater=1979, ic_exrhs=1980,
            ic_xflow=1981, ic_heat=1982, ic_iestim=1983, ic_output=1984, ic_flow=1985,
            ic_return=1986, ic_matfil=1987, ic_sens=1988, ic_sbsout=1991, ic_sbsfil=1992,
            if_idebug=2000, if_wrtinp=2002, if_ncase=2005, if_precision=2009,
            fld_h=1, fld_mu=2, fld_px=3, fld_py=4, fld_pn=5, fld_ux=7, fld_uy=8, fld_un=9,
            fld_taucrt=11, fld_uplsx=12, fld_uplsy=13, fld_sx=15, fld_sy=16, fld_wx=22,
            fld_wy=23, mt_tim=2011, mt_xr=2012, mt_yr=2013, mt_xw=2014, mt_yw=2015, mt_sw=2016,
            mt_run=2021, mt_axle=2022, mt_side=2023,
            err_allow=-12, err_norm=-27, err_tang=-28, err_icp=-31, err_other=-99)

_lib = None
_pi, _pd = POINTER(c_int), POINTER(c_double)

_PROTOS = {
    "cntc_initializefirst": [_pi, _pi, _pi, c_char_p, c_char_p, c_char_p, _pi, _pi, _pi],
    "cntc_initialize": [_pi, _pi, _pi, _pi, c_char_p, _pi],
    "cntc_getmaxnumthreads": [_pi],
    "cntc_setglobalflags": [_pi, _pi, _pi],
    "cntc_setflags": [_pi, _pi, _pi, _pi, _pi],
    "cntc_setsolverflags": [_pi, _pi, _pi, _pi, _pi, _pi, _pd],
    "cntc_setmaterialparameters": [_pi, _pi, _pi, _pi, _pd],
    "cntc_settimestep": [_pi, _pi, _pd],
    "cntc_setreferencevelocity": [_pi, _pi, _pd],
    "cntc_setrollingstepsize": [_pi, _pi, _pd, _pd],
    "cntc_setfrictionmethod": [_pi, _pi, _pi, _pi, _pd],
    "cntc_sethertzcontact": [_pi, _pi, _pi, _pi, _pd],
    "cntc_setpotcontact": [_pi, _pi, _pi, _pi, _pd],
    "cntc_setpenetration": [_pi, _pi, _pd],
    "cntc_setnormalforce": [_pi, _pi, _pd],
    "cntc_setundeformeddistc": [_pi, _pi, _pi, _pi, _pd],
    "cntc_settangentialforces": [_pi, _pi, _pd, _pd],
    "cntc_setcreepages": [_pi, _pi, _pd, _pd, _pd],
    "cntc_setextrarigidslip": [_pi, _pi, _pi, _pd, _pd],
    "cntc_setmetadata": [_pi, _pi, _pi, _pi, _pd],
    "cntc_getflags": [_pi, _pi, _pi, _pi, _pi],
    "cntc_getparameters": [_pi, _pi, _pi, _pi, _pd],
    "cntc_getreferencevelocity": [_pi, _pi, _pd],
    "cntc_getdisplacements": [_pi, _pi, _pi, _pd, _pd, _pd],
    "cntc_calculate": [_pi, _pi, _pi],
    "cntc_calculatebatch": [_pi, _pi, _pi, _pi],
    "cntc_calculatenoalloc": [_pi, _pi, _pi],
    "cntc_servestart": [_pd, _pi],
    "cntc_submit": [_pi, _pi, _pi],
    "cntc_poll": [_pi, _pi, _pi, _pd],
    "cntc_wait": [_pi, _pi, _pd],
    "cntc_servestop": [],
    "cntc_gethertzcontact": [_pi, _pi, _pi, _pd],
    "cntc_getnumelements": [_pi, _pi, _pi, _pi],
    "cntc_getgriddiscretization": [_pi, _pi, _pd, _pd],
    "cntc_getpotcontact": [_pi, _pi, _pi, _pd],
    "cntc_getpenetration": [_pi, _pi, _pd],
    "cntc_getcreepages": [_pi, _pi, _pd, _pd, _pd],
    "cntc_getcontactforces": [_pi, _pi, _pd, _pd, _pd, _pd],
    "cntc_getsensitivities": [_pi, _pi, _pi, _pi, _pd],
    "cntc_getcontactpatchareas": [_pi, _pi, _pd, _pd, _pd],
    "cntc_getelementdivision": [_pi, _pi, _pi, _pi],
    "cntc_getmaximumpressure": [_pi, _pi, _pd],
    "cntc_getmaximumtraction": [_pi, _pi, _pd],
    "cntc_getfielddata": [_pi, _pi, _pi, _pi, _pd],
    "cntc_gettractions": [_pi, _pi, _pi, _pd, _pd, _pd],
    "cntc_getmicroslip": [_pi, _pi, _pi, _pd, _pd],
    "cntc_getiterations": [_pi, _pi, _pi, _pi],
    "cntc_getcalculationtime": [_pi, _pi, _pd, _pd],
    "cntc_resetcalculationtime": [_pi, _pi],
    "cntc_amdprofile": [_pi],
    "cntc_amdprofileget": [_pi, _pi, _pi, _pd, _pd, _pd],
    "cntc_finalize": [_pi],
    "cntc_finalizelast": [],
}


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise RuntimeError("libcntc_amd.so is not built: run contact_amd/build_lib.py")
        L = ctypes.CDLL(library_path)
        for name, args in _PROTOS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        _lib = L
    return _lib


def _i(v):
    return ctypes.byref(c_int(int(v)))


def _d(v):
    return ctypes.byref(c_double(float(v)))


def _iarr(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_pi)


def _darr(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_pd)


def initialize(ire, imodul=3, outdir=" ", idebug=1):
    ifcver, ierr = c_int(0), c_int(0)
    s = outdir.encode()
    lib().cntc_initialize(_i(ire), _i(imodul), ctypes.byref(ifcver), ctypes.byref(ierr), s, _i(len(s)))
    return [ifcver.value, ierr.value]


def getmaxnumthreads():
    v = c_int(0)
    lib().cntc_getmaxnumthreads(ctypes.byref(v))
    return v.value


def setglobalflags(params, values):
    p, pp = _iarr(params)
    v, vp = _iarr(values)
    lib().cntc_setglobalflags(_i(len(p)), pp, vp)


def setflags(ire, icp, params, values):
    p, pp = _iarr(params)
    v, vp = _iarr(values)
    lib().cntc_setflags(_i(ire), _i(icp), _i(len(p)), pp, vp)


def setsolverflags(ire, icp, imeth, iparam, rparam):
    ia, iap = _iarr(iparam)
    ra, rap = _darr(rparam)
    lib().cntc_setsolverflags(_i(ire), _i(icp), _i(imeth), _i(len(ia)), iap, _i(len(ra)), rap)


def setmaterialparameters(ire, icp, m_digit, rparam):
    ra, rap = _darr(rparam)
    lib().cntc_setmaterialparameters(_i(ire), _i(icp), _i(m_digit), _i(len(ra)), rap)


def settimestep(ire, icp, dt):
    lib().cntc_settimestep(_i(ire), _i(icp), _d(dt))


def setreferencevelocity(ire, icp, veloc):
    lib().cntc_setreferencevelocity(_i(ire), _i(icp), _d(veloc))


def setrollingstepsize(ire, icp, chi, dq):
    lib().cntc_setrollingstepsize(_i(ire), _i(icp), _d(chi), _d(dq))


def setfrictionmethod(ire, icp, imeth, params):
    ra, rap = _darr(params)
    lib().cntc_setfrictionmethod(_i(ire), _i(icp), _i(imeth), _i(len(ra)), rap)


def sethertzcontact(ire, icp, ipotcn, params):
    ra, rap = _darr(params)
    lib().cntc_sethertzcontact(_i(ire), _i(icp), _i(ipotcn), _i(len(ra)), rap)


def setpotcontact(ire, icp, ipotcn, params):
    ra, rap = _darr(params)
    lib().cntc_setpotcontact(_i(ire), _i(icp), _i(ipotcn), _i(len(ra)), rap)


def setpenetration(ire, icp, pen):
    lib().cntc_setpenetration(_i(ire), _i(icp), _d(pen))


def setnormalforce(ire, icp, fn):
    lib().cntc_setnormalforce(_i(ire), _i(icp), _d(fn))


def setundeformeddistc(ire, icp, ibase, prmudf):
    ra, rap = _darr(np.ravel(prmudf))
    lib().cntc_setundeformeddistc(_i(ire), _i(icp), _i(ibase), _i(len(ra)), rap)


def settangentialforces(ire, icp, fx, fy):
    """Fx, Fy relative to mu*Fn, for the F digit 1 or 2 (python_intfc/cntc_settangentialforces.py)."""
    lib().cntc_settangentialforces(_i(ire), _i(icp), _d(fx), _d(fy))


def setcreepages(ire, icp, vx, vy, phi):
    lib().cntc_setcreepages(_i(ire), _i(icp), _d(vx), _d(vy), _d(phi))


def setextrarigidslip(ire, icp, wx, wy):
    """Extra term (wx, wy) of the tangential right-hand side per element (E digit 1)."""
    a, ap = _darr(np.ravel(wx))
    b, bp = _darr(np.ravel(wy))
    lib().cntc_setextrarigidslip(_i(ire), _i(icp), _i(min(len(a), len(b))), ap, bp)


def setmetadata(ire, icp, params, values):
    p, pp = _iarr(params)
    v, vp = _darr(values)
    lib().cntc_setmetadata(_i(ire), _i(icp), _i(len(p)), pp, vp)


def getflags(ire, icp, params):
    p, pp = _iarr(params)
    v = np.zeros(len(p), dtype=np.int32)
    lib().cntc_getflags(_i(ire), _i(icp), _i(len(p)), pp, v.ctypes.data_as(_pi))
    return v


def getparameters(ire, icp, itask, lenarr):
    v = np.zeros(lenarr)
    lib().cntc_getparameters(_i(ire), _i(icp), _i(itask), _i(lenarr), v.ctypes.data_as(_pd))
    return v


def getreferencevelocity(ire=1, icp=1):
    v = c_double(0)
    lib().cntc_getreferencevelocity(_i(ire), _i(icp), ctypes.byref(v))
    return v.value


def getdisplacements(ire=1, icp=1):
    """(un, ux, uy) of the solution, each of shape (my, mx)."""
    mx, my = getnumelements(ire, icp)
    un, ux, uy = (np.zeros(mx * my) for _ in range(3))
    lib().cntc_getdisplacements(_i(ire), _i(icp), _i(mx * my), un.ctypes.data_as(_pd),
                                ux.ctypes.data_as(_pd), uy.ctypes.data_as(_pd))
    return un.reshape(my, mx), ux.reshape(my, mx), uy.reshape(my, mx)


def calculate(ire=1, icp=1, idebug=1):
    ierr = c_int(-999)
    lib().cntc_calculate(_i(ire), _i(icp), ctypes.byref(ierr))
    return ierr.value


def calculatenoalloc(ire=1, icp=1):
    """cntc_calculate without device allocation (buffers from the library's fixed array)."""
    ierr = c_int(-999)
    lib().cntc_calculatenoalloc(_i(ire), _i(icp), ctypes.byref(ierr))
    return ierr.value


def calculatebatch(ire_list, icp_list):
    """Solve the cases of several (ire, icp) problems in one batched GPU run."""
    ire, irp = _iarr(ire_list)
    icp, icpp = _iarr(icp_list)
    err = np.zeros(len(ire), dtype=np.int32)
    lib().cntc_calculatebatch(_i(len(ire)), irp, icpp, err.ctypes.data_as(_pi))
    return err


def servestart(budget_bytes=0.0):
    """Start the continuous-batching loop (device-memory budget of active patches)."""
    err = c_int(0)
    lib().cntc_servestart(_d(budget_bytes), ctypes.byref(err))
    return err.value


def submit(ire, icp):
    """Queue the current case of (ire, icp); returns a ticket (> 0) or an error code."""
    t = c_int(0)
    lib().cntc_submit(_i(ire), _i(icp), ctypes.byref(t))
    return t.value


def poll(ticket):
    """(done, ierror, latency_s): done 1 when finished, 0 in flight, -1 unknown ticket."""
    done, err, lat = c_int(0), c_int(0), c_double(0.0)
    lib().cntc_poll(_i(ticket), ctypes.byref(done), ctypes.byref(err), ctypes.byref(lat))
    return done.value, err.value, lat.value


def wait(ticket):
    """Block until the request finishes; returns (ierror, latency_s)."""
    err, lat = c_int(0), c_double(0.0)
    lib().cntc_wait(_i(ticket), ctypes.byref(err), ctypes.byref(lat))
    return err.value, lat.value


def servestop():
    lib().cntc_servestop()


def gethertzcontact(ire=1, icp=1, lenarr=10):
    v = np.zeros(lenarr)
    lib().cntc_gethertzcontact(_i(ire), _i(icp), _i(lenarr), v.ctypes.data_as(_pd))
    return v


def getnumelements(ire=1, icp=1):
    mx, my = c_int(0), c_int(0)
    lib().cntc_getnumelements(_i(ire), _i(icp), ctypes.byref(mx), ctypes.byref(my))
    return mx.value, my.value


def getgriddiscretization(ire=1, icp=1):
    dx, dy = c_double(0), c_double(0)
    lib().cntc_getgriddiscretization(_i(ire), _i(icp), ctypes.byref(dx), ctypes.byref(dy))
    return dx.value, dy.value


def getpenetration(ire=1, icp=1):
    v = c_double(0)
    lib().cntc_getpenetration(_i(ire), _i(icp), ctypes.byref(v))
    return v.value


def getcreepages(ire=1, icp=1):
    """(cksi, ceta, cphi) of the last case: with the F digit 1 or 2 the solved creepages."""
    v = [c_double(0) for _ in range(3)]
    lib().cntc_getcreepages(_i(ire), _i(icp), *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def getcontactforces(ire=1, icp=1):
    v = [c_double(0) for _ in range(4)]
    lib().cntc_getcontactforces(_i(ire), _i(icp), *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def getsensitivities(ire=1, icp=1, lenout=4, lenin=4):
    """Sensitivities [iout, iin] of (Fn, Fx/(mu Fn), Fy/(mu Fn), Mz) to (pen, cksi, ceta, cphi),
    python_intfc/cntc_getsensitivities.py."""
    v = np.zeros(lenout * lenin)
    lib().cntc_getsensitivities(_i(ire), _i(icp), _i(lenout), _i(lenin), v.ctypes.data_as(_pd))
    return v.reshape(lenin, lenout).T.copy()


def getcontactpatchareas(ire=1, icp=1):
    v = [c_double(0) for _ in range(3)]
    lib().cntc_getcontactpatchareas(_i(ire), _i(icp), *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def getelementdivision(ire=1, icp=1):
    mx, my = getnumelements(ire, icp)
    el = np.zeros(mx * my, dtype=np.int32)
    lib().cntc_getelementdivision(_i(ire), _i(icp), _i(mx * my), el.ctypes.data_as(_pi))
    return el.reshape(my, mx)


def getmaximumpressure(ire=1, icp=1):
    v = c_double(0)
    lib().cntc_getmaximumpressure(_i(ire), _i(icp), ctypes.byref(v))
    return v.value


def getfielddata(ire, icp, ifld):
    mx, my = getnumelements(ire, icp)
    v = np.zeros(mx * my)
    lib().cntc_getfielddata(_i(ire), _i(icp), _i(ifld), _i(mx * my), v.ctypes.data_as(_pd))
    return v.reshape(my, mx)


def gettractions(ire=1, icp=1):
    mx, my = getnumelements(ire, icp)
    pn, px, py = (np.zeros(mx * my) for _ in range(3))
    lib().cntc_gettractions(_i(ire), _i(icp), _i(mx * my), pn.ctypes.data_as(_pd),
                            px.ctypes.data_as(_pd), py.ctypes.data_as(_pd))
    return pn.reshape(my, mx), px.reshape(my, mx), py.reshape(my, mx)


def getmicroslip(ire=1, icp=1):
    mx, my = getnumelements(ire, icp)
    sx, sy = np.zeros(mx * my), np.zeros(mx * my)
    lib().cntc_getmicroslip(_i(ire), _i(icp), _i(mx * my), sx.ctypes.data_as(_pd), sy.ctypes.data_as(_pd))
    return sx.reshape(my, mx), sy.reshape(my, mx)


def getiterations(ire=1, icp=1):
    """(itnorm, itcg, ittang, itgs, flags) of the last case of (ire, icp); flags bit 0:
    GDsteady stagnated, bit 1: SteadyGS ran."""
    v = np.zeros(5, dtype=np.int32)
    lib().cntc_getiterations(_i(ire), _i(icp), _i(5), v.ctypes.data_as(_pi))
    return tuple(int(x) for x in v)


def getcalculationtime(ire=1, icp=1):
    c, w = c_double(0), c_double(0)
    lib().cntc_getcalculationtime(_i(ire), _i(icp), ctypes.byref(c), ctypes.byref(w))
    return c.value, w.value


def resetcalculationtime(ire=1, icp=1):
    lib().cntc_resetcalculationtime(_i(ire), _i(icp))


# operation names of the profile keys (key = operation * 4 + stage), in the order of OpKind
OP_NAMES = ["vec", "ncg_negpn", "ncg_dd", "ncg_expand", "ncg_hmin", "sn_contract", "sn_expand",
            "sn_final", "aij_ones", "tcg_nt", "tcg_rset", "tcg_zprep", "tcg_qpost", "tcg_pupd",
            "tcg_as", "tcg_sa", "st_as", "st_sa", "gd_dp", "gd_project", "gd_probe", "gd_update",
            "gd_moves", "gd_diagres", "conv", "conv32", "elascf", "embed", "extract", "setup",
            "forces", "stdygs", "guard", "sens", "sens_ext", "ledge", "gsb_setup", "gsb_walk",
            "gsb_check", "cnvxgs", "dense", "tcg_vupd"]


def amdprofile(on):
    """Start (True: every kernel, or 2: only the SteadyGS, GD probe, FFT and dense product kernels;
    clears earlier records) or stop (False) per-kernel HIP-event timing."""
    lib().cntc_amdprofile(_i(2 if on == 2 else (1 if on else 0)))


def amdprofileget(lenarr=256):
    """Per-kernel records since amdprofile(True): {name: (seconds, algorithmic bytes, launches)};
    for the SteadyGS kernels the second value is algorithmic flops."""
    kinds = np.zeros(lenarr, dtype=np.int32)
    sec, byt, nl = (np.zeros(lenarr) for _ in range(3))
    n = c_int(0)
    lib().cntc_amdprofileget(_i(lenarr), ctypes.byref(n), _iarr(kinds)[1], _darr(sec)[1],
                             _darr(byt)[1], _darr(nl)[1])
    out = {}
    for k in range(min(n.value, lenarr)):
        op, stage = divmod(int(kinds[k]), 4)
        name = OP_NAMES[op] if op < len(OP_NAMES) else "op%d" % op
        if op in (24, 25):
            name += ("_fwd", "_mul", "_inv")[stage]
        elif name == "stdygs" and stage == 1:
            name = "stdygs_staged"
        elif name == "stdygs" and stage == 2:
            name = "stdygs_overlap"
        elif name == "stdygs" and stage == 3:
            name = "stdygs_block"
        out[name] = (float(sec[k]), float(byt[k]), int(nl[k]))
    return out


def finalize(ire):
    lib().cntc_finalize(_i(ire))


def finalizelast():
    lib().cntc_finalizelast()
