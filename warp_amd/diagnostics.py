"""Whole-beam moments and emittances, as the reference's getzmmnt
(pywarp90/top.F:521-1500) leaves them in top: the 33 moment sums zmmnts0
are accumulated on the GPU over every particle of the selected species
(libwarpmi ``wm_getzmmnt``), then finalized here in FP64 with the
reference's formulas (top.F:1363-1500): means, rms sizes and velocities,
the transverse, longitudinal and generalized emittances.  Moments of all
species are combined (lspeciesmoments false, the reference's default).
"""
import ctypes as C
import math

import torch

from . import _native as N
from .globals import clight

SMALLPOS = 1.0e-36  # pywarp90/top.h:4

_NAMES = ("pnum", "xbar", "ybar", "zbar", "xpbar", "ypbar", "vxbar", "vybar", "vzbar",
          "xybar", "xpypbar", "xsqbar", "ysqbar", "zsqbar", "xpsqbar", "ypsqbar", "vxsqbar",
          "vysqbar", "vzsqbar", "xxpbar", "yypbar", "zvzbar", "xypbar", "yxpbar", "xvzbar",
          "yvzbar", "vxvzbar", "vyvzbar", "xvxbar", "yvybar", "xvybar", "yvxbar", "vxvybar")


def _dvnz(x):
    """dvnz(X) = sign(abs(X)+SMALLPOS, X) (top.h:27)."""
    return math.copysign(abs(x) + SMALLPOS, x)


def zmmnts0(pgroup, jslist=None, lrelativ=False, l_momentum_moments=False):
    """The moment sums zmmnts0(1:33) over the species in jslist (all by
    default), as a list of 33 floats (index k-1 for zmmnts0(k))."""
    usegaminv = int(lrelativ and not l_momentum_moments)
    total = [0.] * 33
    dev = pgroup.device
    if dev.type != "cuda":
        raise RuntimeError("getzmmnt runs on the HIP device")
    s = N.stream_of(dev)
    mm = torch.zeros(33, dtype=torch.float64, device=dev)
    for js in (range(pgroup.ns) if jslist is None else jslist):
        sl = pgroup.species_slice(js)
        n = sl.stop - sl.start
        if n == 0:
            continue
        work = torch.empty(int(N.lib().wm_getzmmnt_worksize(n)), dtype=torch.uint8, device=dev)
        col = lambda c: getattr(pgroup, c)[sl].data_ptr()
        N.call("wm_getzmmnt", n, col("xp"), col("yp"), col("zp"), col("uxp"), col("uyp"),
               col("uzp"), col("gaminv"), usegaminv, mm.data_ptr(), work.data_ptr(), s)
        part = mm.cpu().tolist()
        total = [a + b for a, b in zip(total, part)]
    return total


def finalize(mm):
    """top.F:1363-1480 for one moment set: returns a dict with the
    reference's names (pnum, xbar, ..., xrms, ..., epsx, epsnx, ..., epsh,
    epsnh)."""
    clighti = 1. / clight
    pnumi = 1. / (mm[0] + SMALLPOS)
    d = {"pnum": mm[0]}
    for k in range(1, 33):
        d[_NAMES[k]] = mm[k] * pnumi
    g = d.get
    delxsq = g("xsqbar") - g("xbar") * g("xbar")
    delxxp = g("xxpbar") - g("xbar") * g("xpbar")
    delxpsq = g("xpsqbar") - g("xpbar") * g("xpbar")
    delxvx = g("xvxbar") - g("xbar") * g("vxbar")
    delvxsq = g("vxsqbar") - g("vxbar") ** 2
    delysq = g("ysqbar") - g("ybar") * g("ybar")
    delyyp = g("yypbar") - g("ybar") * g("ypbar")
    delypsq = g("ypsqbar") - g("ypbar") * g("ypbar")
    delyvy = g("yvybar") - g("ybar") * g("vybar")
    delvysq = g("vysqbar") - g("vybar") ** 2
    delxy = g("xybar") - g("xbar") * g("ybar")
    delxyp = g("xypbar") - g("xbar") * g("ypbar")
    delyxp = g("yxpbar") - g("ybar") * g("xpbar")
    delxpyp = g("xpypbar") - g("xpbar") * g("ypbar")
    delxvy = g("xvybar") - g("xbar") * g("vybar")
    delyvx = g("yvxbar") - g("ybar") * g("vxbar")
    delvxvy = g("vxvybar") - g("vxbar") * g("vybar")
    delzsq = g("zsqbar") - g("zbar") ** 2
    delvzsq = g("vzsqbar") - g("vzbar") ** 2
    delzvz = g("zvzbar") - g("zbar") * g("vzbar")
    r = math.sqrt
    d["xrms"] = r(max(0., delxsq))
    d["yrms"] = r(max(0., delysq))
    d["zrms"] = r(max(0., delzsq))
    d["rrms"] = r(d["xrms"] ** 2 + d["yrms"] ** 2)
    d["xprms"] = r(max(0., delxpsq))
    d["yprms"] = r(max(0., delypsq))
    d["vxrms"] = r(max(0., delvxsq))
    d["vyrms"] = r(max(0., delvysq))
    d["vzrms"] = r(max(0., delvzsq))
    d["epsx"] = 4. * r(max(0., delxsq * delxpsq - delxxp ** 2))
    d["epsy"] = 4. * r(max(0., delysq * delypsq - delyyp ** 2))
    d["epsnx"] = 4. * r(max(0., delxsq * delvxsq - delxvx ** 2)) * clighti * 1.e6
    d["epsny"] = 4. * r(max(0., delysq * delvysq - delyvy ** 2)) * clighti * 1.e6
    d["epsnz"] = 4. * r(max(0., delzsq * delvzsq - delzvz ** 2)) * clighti
    d["epsz"] = d["epsnz"] / _dvnz(g("vzbar") * clighti)
    tr = 4. * ((delxsq + delysq) * (delxpsq + delypsq)
               - (delxxp + delyyp) ** 2 - (delxyp - delyxp) ** 2)
    d["epsr"] = r(max(0., tr))
    tg = 0.5 * (d["epsx"] ** 2 + d["epsy"] ** 2) + 16 * (delxy * delxpyp - delxyp * delyxp)
    d["epsg"] = r(max(0., tg))
    th = (d["epsx"] ** 2 * d["epsy"] ** 2 +
          256 * ((delxy * delxpyp) ** 2 + (delxyp * delyxp) ** 2 -
                 delxsq * delysq * (delxpyp) ** 2 - delxsq * delypsq * (delyxp) ** 2 -
                 delxpsq * delysq * (delxyp) ** 2 - delxpsq * delypsq * (delxy) ** 2 -
                 2 * delxy * delxyp * delyxp * delxpyp + 2 * delxxp * delypsq * delxy * delyxp -
                 2 * delxxp * delyyp * delxy * delxpyp - 2 * delxxp * delyyp * delxyp * delyxp +
                 2 * delxpsq * delyyp * delxy * delxyp + 2 * delxsq * delyyp * delyxp * delxpyp +
                 2 * delxxp * delysq * delxpyp * delxyp))
    d["epsh"] = r(r(max(0., th)))
    tr = 4. * ((delxsq + delysq) * (delvxsq + delvysq)
               - (delxvx + delyvy) ** 2 - (delxvy - delyvx) ** 2)
    d["epsnr"] = r(max(0., tr)) * clighti * 1.e6
    tg = (0.5 * (d["epsnx"] ** 2 + d["epsny"] ** 2) +
          16 * (delxy * delvxvy - delxvy * delyvx) * clighti ** 2 * 1.e12)
    d["epsng"] = r(max(0., tg))
    th = (d["epsnx"] ** 2 * d["epsny"] ** 2 +
          256 * ((delxy * delvxvy) ** 2 + (delxvy * delyvx) ** 2 -
                 delxsq * delysq * (delvxvy) ** 2 - delxsq * delvysq * (delyvx) ** 2 -
                 delvxsq * delysq * (delxvy) ** 2 - delvxsq * delvysq * (delxy) ** 2 -
                 2 * delxy * delxvy * delyvx * delvxvy + 2 * delxvx * delvysq * delxy * delyvx -
                 2 * delxvx * delyvy * delxy * delvxvy - 2 * delxvx * delyvy * delxvy * delyvx +
                 2 * delvxsq * delyvy * delxy * delxvy + 2 * delxsq * delyvy * delyvx * delvxvy +
                 2 * delxvx * delysq * delvxvy * delxvy) * clighti ** 4 * 1.e24)
    d["epsnh"] = r(r(max(0., th)))
    d["bmlen"] = r(max(0., g("zsqbar") - g("zbar") ** 2))
    return d


def getzmmnt(pgroup, jslist=None, lrelativ=False, l_momentum_moments=False):
    """Moments and emittances of the beam (getzmmnt itask 1-3, whole-beam
    part), as a dict keyed by the reference's top variable names."""
    return finalize(zmmnts0(pgroup, jslist, lrelativ, l_momentum_moments))
