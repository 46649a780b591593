"""ctypes binding of the CPU restatement in cntc_oracle.c.

Test infrastructure only: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product (contact_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libcntc_oracle.so")

EXTER, ADHES, SLIP, PLAST = 0, 1, 2, 3


class OrcInput(ctypes.Structure):
    _fields_ = [
        ("tang", ctypes.c_int), ("norm", ctypes.c_int), ("pvtime", ctypes.c_int),
        ("iestim", ctypes.c_int), ("gausei", ctypes.c_int),
        ("poiss1", ctypes.c_double), ("poiss2", ctypes.c_double),
        ("gg1", ctypes.c_double), ("gg2", ctypes.c_double),
        ("fstat", ctypes.c_double),
        ("ipotcn", ctypes.c_int), ("mx", ctypes.c_int), ("my", ctypes.c_int),
        ("xc1", ctypes.c_double), ("yc1", ctypes.c_double),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("hz_a1", ctypes.c_double), ("hz_b1", ctypes.c_double),
        ("hz_aa", ctypes.c_double), ("hz_bb", ctypes.c_double), ("hz_scale", ctypes.c_double),
        ("hz_aob", ctypes.c_double),
        ("ibase", ctypes.c_int), ("prmudf", ctypes.c_double * 6),
        ("h_elem", ctypes.POINTER(ctypes.c_double)),
        ("pen", ctypes.c_double), ("fn", ctypes.c_double), ("cksi", ctypes.c_double),
        ("ceta", ctypes.c_double), ("cphi", ctypes.c_double), ("chi", ctypes.c_double),
        ("dq", ctypes.c_double), ("veloc", ctypes.c_double),
        ("maxgs", ctypes.c_int), ("maxin", ctypes.c_int), ("maxnr", ctypes.c_int),
        ("maxout", ctypes.c_int), ("eps", ctypes.c_double),
        ("fdecay", ctypes.c_double), ("betath", ctypes.c_double), ("d_ifc", ctypes.c_double),
        ("d_lin", ctypes.c_double), ("d_cns", ctypes.c_double), ("d_slp", ctypes.c_double),
        ("pow_s", ctypes.c_double), ("kdowfb", ctypes.c_int),
        ("omegah", ctypes.c_double), ("omegas", ctypes.c_double),
        ("sens", ctypes.c_int), ("mxsens", ctypes.c_int), ("epsens", ctypes.c_double),
        ("force3", ctypes.c_int), ("fxrel", ctypes.c_double), ("fyrel", ctypes.c_double),
        ("rztang", ctypes.c_int), ("exrhs", ctypes.POINTER(ctypes.c_double)),
        ("bound", ctypes.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.orc_new.restype = vp
        L.orc_free.argtypes = [vp]
        L.orc_default_input.argtypes = [ctypes.POINTER(OrcInput)]
        L.orc_calculate.argtypes = [vp, ctypes.POINTER(OrcInput)]
        L.orc_calculate.restype = ctypes.c_int
        L.orc_get_mx.argtypes = [vp]
        L.orc_get_my.argtypes = [vp]
        L.orc_get_grid.argtypes = [vp, dp, dp, dp, dp]
        L.orc_get_forces.argtypes = [vp, dp, dp, dp, dp]
        L.orc_get_tractions.argtypes = [vp, dp, dp, dp]
        L.orc_get_slip.argtypes = [vp, dp, dp]
        L.orc_get_eldiv.argtypes = [vp, ip]
        L.orc_get_h.argtypes = [vp, dp]
        L.orc_get_pen.argtypes = [vp]
        L.orc_get_pen.restype = ctypes.c_double
        L.orc_get_hertz.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.orc_get_displacements.argtypes = [vp, dp, dp, dp]
        L.orc_get_iters.argtypes = [vp, ip, ip, ip, ip]
        L.orc_get_stagnated.argtypes = [vp]
        L.orc_get_stagnated.restype = ctypes.c_int
        L.orc_get_sens.argtypes = [vp, dp]
        L.orc_get_creepages.argtypes = [vp, dp, dp, dp]
        L.orc_linrol.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.orc_opt_fft_size.argtypes = [ctypes.c_int]
        L.orc_opt_fft_size.restype = ctypes.c_int
        L.orc_elascf_pcwcns.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, dp]
        L.orc_conv_block.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, ctypes.c_int]
        L.orc_make_prec_block.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp]
        L.orc_ellip_bd.argtypes = [ctypes.c_double, dp, dp]
        L.orc_fft.argtypes = [ctypes.c_int, dp, dp, dp, dp, ctypes.c_int]
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_input():
    inp = OrcInput()
    lib().orc_default_input(ctypes.byref(inp))
    return inp


class Problem:
    """One (ire, icp) contact problem solved by the CPU restatement."""

    def __init__(self):
        self._p = lib().orc_new()
        self._h = None

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_free(self._p)
            self._p = None

    def calculate(self, inp, h_elem=None):
        if h_elem is not None:
            self._h = np.ascontiguousarray(h_elem, dtype=np.float64)
            inp.h_elem = _dptr(self._h)
        return lib().orc_calculate(self._p, ctypes.byref(inp))

    @property
    def shape(self):
        return lib().orc_get_mx(self._p), lib().orc_get_my(self._p)

    def forces(self):
        v = [ctypes.c_double() for _ in range(4)]
        lib().orc_get_forces(self._p, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def tractions(self):
        mx, my = self.shape
        pn, px, py = (np.zeros(mx * my) for _ in range(3))
        lib().orc_get_tractions(self._p, _dptr(pn), _dptr(px), _dptr(py))
        return pn, px, py

    def slip(self):
        mx, my = self.shape
        sx, sy = np.zeros(mx * my), np.zeros(mx * my)
        lib().orc_get_slip(self._p, _dptr(sx), _dptr(sy))
        return sx, sy

    def eldiv(self):
        mx, my = self.shape
        el = np.zeros(mx * my, dtype=np.int32)
        lib().orc_get_eldiv(self._p, el.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return el

    def h(self):
        mx, my = self.shape
        h = np.zeros(mx * my)
        lib().orc_get_h(self._p, _dptr(h))
        return h

    def pen(self):
        return lib().orc_get_pen(self._p)

    def grid(self):
        v = [ctypes.c_double() for _ in range(4)]
        lib().orc_get_grid(self._p, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def displacements(self):
        """(un, ux, uy) = A ps on the interior, 0 elsewhere (flat, x fastest)."""
        mx, my = self.shape
        un, ux, uy = (np.zeros(mx * my) for _ in range(3))
        lib().orc_get_displacements(self._p, _dptr(un), _dptr(ux), _dptr(uy))
        return un, ux, uy

    def hertz(self):
        v = [ctypes.c_double() for _ in range(6)]
        lib().orc_get_hertz(self._p, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def stagnated(self):
        return lib().orc_get_stagnated(self._p)

    def sens(self):
        """4x4 sensitivities in internal units, [iout, iin]: outputs (fn, fx/(mu fn),
        fy/(mu fn), mz), inputs (pen, cksi, ceta, cphi)."""
        s = np.zeros(16)
        lib().orc_get_sens(self._p, _dptr(s))
        return s.reshape(4, 4).T.copy()

    def creepages(self):
        v = [ctypes.c_double() for _ in range(3)]
        lib().orc_get_creepages(self._p, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def iters(self):
        v = [ctypes.c_int() for _ in range(4)]
        lib().orc_get_iters(self._p, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)


def linrol(nu, aob):
    v = [ctypes.c_double() for _ in range(3)]
    lib().orc_linrol(nu, aob, *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def opt_fft_size(n):
    return lib().orc_opt_fft_size(int(n))


def elascf_pcwcns(akv, nuv, mx, my, dx, dy, xshft=0.0, yshft=0.0):
    cf = np.zeros(9 * (2 * mx) * (2 * my))
    lib().orc_elascf_pcwcns(akv, nuv, mx, my, dx, dy, xshft, yshft, _dptr(cf))
    return cf.reshape(3, 3, 2 * my, 2 * mx)


def conv_block(cfblock, p, method=1):
    my2, mx2 = cfblock.shape
    mx, my = mx2 // 2, my2 // 2
    cfb = np.ascontiguousarray(cfblock, dtype=np.float64)
    pp = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
    u = np.zeros(mx * my)
    lib().orc_conv_block(mx, my, _dptr(cfb), _dptr(pp), _dptr(u), method)
    return u


def make_prec_block(cfblock, ga):
    my2, mx2 = cfblock.shape
    cfb = np.ascontiguousarray(cfblock, dtype=np.float64)
    m = np.zeros_like(cfb)
    lib().orc_make_prec_block(mx2 // 2, my2 // 2, ga, _dptr(cfb), _dptr(m))
    return m


def ellip_bd(mc):
    b, d = ctypes.c_double(), ctypes.c_double()
    lib().orc_ellip_bd(mc, ctypes.byref(b), ctypes.byref(d))
    return b.value, d.value


def fft(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    re, im = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    ro, io = np.zeros(len(x)), np.zeros(len(x))
    lib().orc_fft(len(x), _dptr(re), _dptr(im), _dptr(ro), _dptr(io), sign)
    return ro + 1j * io


# GDsteady parameters of the reference's steady-rolling inputs (perfc_test/tang_problm_1c.inp:12)
GD_PARAMS = dict(fdecay=1.0, betath=0.05, kdowfb=1, d_ifc=2.0, d_lin=-1.0, d_cns=1.0, d_slp=2.6,
                 pow_s=1.0)


def case_input(c):
    """Oracle input record (and element-wise gap or None) of a case dict as issued by
    contact_amd.workloads.apply_case."""
    inp = default_input()
    inp.tang = c["tang"]
    inp.norm = c["norm"]
    inp.fn = c.get("fn", 0.0)
    inp.pen = c.get("pen", 0.0)
    inp.poiss1, inp.poiss2 = c["nu"], c.get("nu2", c["nu"])
    inp.gg1, inp.gg2 = c["g"], c.get("g2", c["g"])
    inp.fstat = c["fstat"]
    inp.veloc = c.get("veloc", 10000.0)
    inp.maxgs = c["maxgs"]
    inp.maxin = c["maxin"]
    inp.maxnr = 30
    inp.maxout = 1
    inp.eps = c["eps"]
    g = GD_PARAMS
    inp.gausei = c.get("gausei", 5)
    inp.fdecay, inp.betath, inp.kdowfb = g["fdecay"], g["betath"], g["kdowfb"]
    inp.d_ifc, inp.d_lin, inp.d_cns, inp.d_slp, inp.pow_s = (g["d_ifc"], g["d_lin"], g["d_cns"],
                                                             g["d_slp"], g["pow_s"])
    if inp.gausei in (2, 3):
        inp.omegah, inp.omegas = c.get("omegah", 0.5), c.get("omegas", 0.5)
    if c.get("pvtime") is not None:
        inp.pvtime = c["pvtime"]
        inp.iestim = c.get("iestim", 0)
    h = None
    if c["ipotcn"] == -3:
        inp.ipotcn = -3
        inp.mx, inp.my = c["mx"], c["my"]
        inp.hz_aa, inp.hz_bb, inp.hz_scale = c["aa"], c["bb"], c["scale"]
    elif c["ipotcn"] == -2:
        inp.ipotcn = -2
        inp.mx, inp.my = c["mx"], c["my"]
        inp.hz_a1, inp.hz_aob, inp.hz_scale = c["a1"], c["aob"], c["scale"]
    elif c["ipotcn"] == -1:
        inp.ipotcn = -1
        inp.mx, inp.my = c["mx"], c["my"]
        inp.hz_a1, inp.hz_b1, inp.hz_scale = c["a1"], c["b1"], c["scale"]
    elif c["ipotcn"] in (-4, -5):
        inp.ipotcn = c["ipotcn"]
        inp.mx, inp.my = c["mx"], c["my"]
        if c["ipotcn"] == -4:
            inp.hz_a1 = c["a1"]
        else:
            inp.hz_aa = c["aa"]
        inp.hz_bb, inp.hz_scale = c["bb"], c["scale"]
        inp.bound = 2
    else:
        inp.ipotcn = 3
        inp.mx, inp.my = c["mx"], c["my"]
        inp.dx, inp.dy = c["dx"], c["dy"]
        inp.xc1 = c["xl"] + 0.5 * c["dx"]
        inp.yc1 = c["yl"] + 0.5 * c["dy"]
        if c.get("h") is not None:
            inp.ibase = 9
            h = np.asarray(c["h"], dtype=np.float64)
        else:
            inp.ibase = 1
            for k in range(6):
                inp.prmudf[k] = c["prmudf"][k]
    if c.get("dq") is not None:
        inp.dq = c["dq"]
    inp.cksi, inp.ceta, inp.cphi = c["cksi"], c["ceta"], c["cphi"]
    inp.force3 = c.get("force", 0)
    fx, fy = c.get("fxrel", 0.0), c.get("fyrel", 0.0)
    fabs = (fx * fx + fy * fy) ** 0.5
    if fabs > 1.0:      # cntc_setTangentialForces normalises |(fx, fy)| to 1
        fx, fy = fx / fabs, fy / fabs
    inp.fxrel, inp.fyrel = fx, fy
    if c.get("rztang") is not None:
        inp.rztang = c["rztang"]
    if c.get("bound") is not None:
        inp.bound = c["bound"]
    if c.get("exrhs") is not None:
        # extra rigid slip, x then y (cntc_setExtraRigidSlip); kept alive with the record
        inp.rztang = 1
        ex = np.ascontiguousarray(np.concatenate([np.ravel(c["exrhs"][0]), np.ravel(c["exrhs"][1])]),
                                  dtype=np.float64)
        inp._exrhs_keep = ex
        inp.exrhs = _dptr(ex)
    inp.sens = c.get("sens", 1)
    inp.mxsens = c.get("mxsens", 30)
    inp.epsens = c.get("epsens", 1e-4)
    return inp, h


def solve_case(c):
    """Solve a case dict with the CPU restatement; returns (rc, forces, problem)."""
    P = Problem()
    inp, h = case_input(c)
    rc = P.calculate(inp, h)
    return rc, P.forces(), P
