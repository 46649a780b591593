// inp_writer.cpp -- the .inp description of a module-3 case (wrtinp, src/m_sinput.f90:1506-1862)
#include "inp_writer.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace cntc {

static std::string pad_left(const std::string& s, int w)
{
    return (int)s.size() >= w ? s : std::string(w - s.size(), ' ') + s;
}

std::string fort_i(long v, int w, int m)
{
    char b[64];
    if (m > 0) std::snprintf(b, sizeof b, "%*.*ld", w, m, v);
    else std::snprintf(b, sizeof b, "%*ld", w, v);
    return b;
}

std::string fort_f(double x, int w, int d)
{
    char b[96];
    std::snprintf(b, sizeof b, "%*.*f", w, d, x);
    return b;
}

// Ew.d: sign, "0.", d digits, "E", exponent sign and two digits
std::string fort_e(double x, int w, int d)
{
    if (x == 0.0) {
        std::string s = "0." + std::string(d, '0') + "E+00";
        return pad_left(s, w);
    }
    char b[96];
    std::snprintf(b, sizeof b, "%.*e", d - 1, std::fabs(x));   // m.mmm e+XX
    const char* e = std::strchr(b, 'e');
    const int ex = std::atoi(e + 1) + 1;
    std::string dig;
    for (const char* p = b; p < e; p++)
        if (*p != '.') dig += *p;
    char eb[16];
    std::snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
    return pad_left(std::string(x < 0 ? "-" : "") + "0." + dig + eb, w);
}

// ESw.d: one nonzero digit before the point
std::string fort_es(double x, int w, int d)
{
    char b[96];
    std::snprintf(b, sizeof b, "%*.*E", w, d, x);
    return b;
}

// Gw.d: F form with d significant digits and four trailing blanks when 0.1 <= |x| (rounded to
// d digits) < 10^d, or zero; E form otherwise
std::string fort_g(double x, int w, int d)
{
    if (x == 0.0) return fort_f(0.0, w - 4, d - 1) + "    ";
    char b[96];
    std::snprintf(b, sizeof b, "%.*e", d - 1, std::fabs(x));
    const int ex = std::atoi(std::strchr(b, 'e') + 1);
    if (ex >= -1 && ex <= d - 1) return fort_f(x, w - 4, d - (ex + 1)) + "    ";
    return fort_e(x, w, d);
}

// format 7509: ( 5g14.6,:,/, 9(2x,5g14.6,:,/), 10( /, 10(2x,5g14.6,:,/) ) ), with reversion to
// the last group for longer rows
static void write_row_7509(std::FILE* f, const double* v, int n)
{
    int k = 0;
    std::string line;
    for (; k < std::min(n, 5); k++) line += fort_g(v[k], 14, 6);
    std::fprintf(f, "%s\n", line.c_str());
    for (int l = 0; l < 9 && k < n; l++) {
        line = "  ";
        for (int j = 0; j < 5 && k < n; j++, k++) line += fort_g(v[k], 14, 6);
        std::fprintf(f, "%s\n", line.c_str());
    }
    while (k < n) {
        std::fprintf(f, "\n");
        for (int l = 0; l < 10 && k < n; l++) {
            line = "  ";
            for (int j = 0; j < 5 && k < n; j++, k++) line += fort_g(v[k], 14, 6);
            std::fprintf(f, "%s\n", line.c_str());
        }
    }
}

void write_inp_case(std::FILE* f, int ire, int icp, int ncase, const Patch& P, const InpDigits& dg,
                    double fkin, const std::vector<int8_t>& el)
{
    const CaseInput& in = P.in;
    const int mx = in.mx, my = in.my;
    const bool is_roll = in.tang == 2 || in.tang == 3;
    // header of the case (src/contact_addon.f90:3862-3863)
    std::fprintf(f, "%s%s%s  MODULE, RE, CP: case %d for result element%s, contact problem%s\n",
                 fort_i(3, 4).c_str(), fort_i(ire, 4).c_str(), fort_i(icp, 4).c_str(), ncase,
                 fort_i(ire, 3).c_str(), fort_i(icp, 3).c_str());

    // the box around the contact area (ltight, areas: src/m_gridfunc.f90:582-637); without a
    // solution the contact area is empty
    int ix0 = 1, ix1 = mx, iy0 = 1, iy1 = my;
    if (mx > 0 && my > 0) {
        const bool have = (size_t)mx * my == el.size();
        int ixmin = mx, ixmax = 1, iymin = my, iymax = 1;
        for (int iy = 1; have && iy <= my; iy++) {
            const int8_t* row = el.data() + (size_t)(iy - 1) * mx;
            int first = 1;
            while (row[first - 1] <= 0 && first != mx) first++;
            int last = 0;
            for (int ix = first; ix <= mx; ix++)
                if (row[ix - 1] >= 1) last = ix;
            if (first <= last) {
                ixmin = std::min(ixmin, first);
                ixmax = std::max(ixmax, last);
                iymin = std::min(iymin, iy);
                iymax = std::max(iymax, iy);
            }
        }
        if (ixmin > ixmax) {
            ix0 = std::max(1, (mx + 1) / 2 - 2);
            ix1 = std::min(mx, (mx + 1) / 2 + 2);
            iy0 = std::max(1, (my + 1) / 2 - 2);
            iy1 = std::min(my, (my + 1) / 2 + 2);
        } else {
            ix0 = std::max(1, ixmin - 2);
            ix1 = std::min(mx, ixmax + 2);
            iy0 = std::max(1, iymin - 2);
            iy1 = std::min(my, iymax + 2);
        }
    }

    // control digits (ic_pack, src/m_hierarch_data.f90:1220-1252; D = C = Z = 2: a new grid,
    // new influence coefficients and a new undeformed distance in every case written)
    const long pbtnfs = 100000L * in.pvtime + 10000L * in.bound + 1000L * in.tang + 100L * in.norm +
                        10L * in.force3 + dg.stress;
    const long vldcmze = 100000L * dg.frclaw + 10000L * 2 + 1000L * 2 + 100L * dg.mater + 10L * 2 + in.rztang;
    const long xgiaowr = 100000L * in.gausei + 10000L * in.iestim + 1000L * dg.matfil + 100L * dg.output +
                         10L * dg.flow + dg.ret;
    std::fprintf(f, "%s      %s\n", fort_i(pbtnfs, 8, 6).c_str(),
                 "    P-B-T-N-F-S           PVTIME, BOUND,  TANG,   NORM,   FORCE,  STRESS");
    std::fprintf(f, "%s      %s\n", fort_i(vldcmze, 8, 6).c_str(),
                 "    L-D-C-M-Z-E           FRCLAW, DISCNS, INFLCF, MATER,  RZNORM, EXRHS");
    std::fprintf(f, "%s      %s\n", fort_i(xgiaowr, 8, 7).c_str(),
                 "  H-G-I-A-O-W-R     HEAT, GAUSEI, IESTIM, MATFIL, OUTPUT, FLOW,   RETURN");

    // solver settings
    if (in.gausei != 1)
        std::fprintf(f, "%s%s%s%s   %s        MAXGS , MAXIN , MAXNR , MAXOUT, EPS\n",
                     fort_i(in.maxgs, 6).c_str(), fort_i(in.maxin, 6).c_str(), fort_i(in.maxnr, 5).c_str(),
                     fort_i(in.maxout, 5).c_str(), fort_es(in.eps, 8, 1).c_str());
    if (in.gausei == 2 || in.gausei == 3)
        std::fprintf(f, "%s%s%s%s       OMEGAH, OMEGAS, INISLP, OMGSLP\n", fort_g(in.omegah, 12, 4).c_str(),
                     fort_g(in.omegas, 12, 4).c_str(), fort_i(0, 6).c_str(), fort_g(0.5, 12, 4).c_str());
    if (in.gausei == 4)
        std::fprintf(f, "%s%s%40sINISLP, OMGSLP\n", fort_i(0, 6).c_str(), fort_g(0.5, 12, 4).c_str(), "");
    if (in.gausei == 5) {
        // the eight values on one line, KDOWFB as an integer: what the reader takes
        // ('ddiddddd', src/m_sinput.f90:648-658)
        std::fprintf(f, "%s%s%s%s%s%s%s%s          FDECAY, D_IFC/LIN/CNS, D_SLP, POW_S\n",
                     fort_f(in.fdecay, 8, 3).c_str(), fort_f(in.betath, 8, 3).c_str(),
                     fort_i(in.kdowfb, 8).c_str(), fort_f(in.d_ifc, 8, 3).c_str(),
                     fort_f(in.d_lin, 8, 3).c_str(), fort_f(in.d_cns, 8, 3).c_str(),
                     fort_f(in.d_slp, 8, 3).c_str(), fort_f(in.pow_s, 8, 3).c_str());
    }

    // kinematic constants (no spin of the bodies)
    const double fun = in.norm == 0 ? in.pen : in.fn;
    const double fux = in.force3 == 0 ? in.cksi : in.fxrel;
    const double fuy = in.force3 == 2 ? in.fyrel : in.ceta;
    static const char* kin_lbl[2][3] = {{"PEN, CKSI, CETA, CPHI", "PEN, FX, CETA, CPHI", "PEN, FX, FY, CPHI"},
                                        {"FN, CKSI, CETA, CPHI", "FN, FX, CETA, CPHI", "FN, FX, FY, CPHI"}};
    std::fprintf(f, "%s%s%s%s          %s\n", fort_g(fun, 12, 4).c_str(), fort_g(fux, 12, 4).c_str(),
                 fort_g(fuy, 12, 4).c_str(), fort_g(in.cphi, 12, 4).c_str(),
                 kin_lbl[in.norm ? 1 : 0][std::min(2, std::max(0, in.force3))]);

    // friction (L = 0: Coulomb, fric_wrtinp, src/m_friclaw.f90:869-905)
    std::fprintf(f, "%s%s%34sFSTAT, FKIN\n", fort_g(in.fstat, 12, 4).c_str(), fort_g(fkin, 12, 4).c_str(), "");

    // rolling step and material (C = 2)
    if (is_roll) {
        const double dq = in.dq > 0.0 ? in.dq : P.dx;
        std::fprintf(f, "%s%s%s%22sCHI, DQ, VELOC\n", fort_g(in.chi, 12, 4).c_str(), fort_g(dq, 12, 4).c_str(),
                     fort_g(in.veloc, 12, 4).c_str(), "");
    }
    std::fprintf(f, "%s    %s    %s%s          POISS 1,2,  GG 1,2\n", fort_f(in.poiss1, 8, 4).c_str(),
                 fort_f(in.poiss2, 8, 4).c_str(), fort_g(in.gg1, 12, 4).c_str(), fort_g(in.gg2, 12, 4).c_str());

    // discretisation (D = 2)
    std::fprintf(f, "%s%53sIPOTCN\n", fort_i(in.ipotcn, 5).c_str(), "");
    const std::string mxy = fort_i(mx, 5) + fort_i(my, 5);
    switch (in.ipotcn) {
    case -1:
        std::fprintf(f, "%s%s%s%s            MX,MY,A1,B1,SCALE\n", mxy.c_str(), fort_g(in.hz_a1, 12, 4).c_str(),
                     fort_g(in.hz_b1, 12, 4).c_str(), fort_g(in.hz_scale, 12, 4).c_str());
        break;
    case -2:
        std::fprintf(f, "%s%s%s%s            MX,MY,A1,AOB,SCALE\n", mxy.c_str(), fort_g(in.hz_a1, 12, 4).c_str(),
                     fort_g(in.hz_aob, 12, 4).c_str(), fort_g(in.hz_scale, 12, 4).c_str());
        break;
    case -3: case -5:
        std::fprintf(f, "%s%s%s%s            MX,MY,AA,BB,SCALE\n", mxy.c_str(), fort_g(in.hz_aa, 12, 4).c_str(),
                     fort_g(in.hz_bb, 12, 4).c_str(), fort_g(in.hz_scale, 12, 4).c_str());
        break;
    case -4:
        std::fprintf(f, "%s%s%s%s            MX,MY,A1,BB,SCALE\n", mxy.c_str(), fort_g(in.hz_a1, 12, 4).c_str(),
                     fort_g(in.hz_bb, 12, 4).c_str(), fort_g(in.hz_scale, 12, 4).c_str());
        break;
    default: {
        // grid forms 1-4, clipped to the box; every form is written with label 6301
        const double dx = in.dx, dy = in.dy;
        double a = 0, b = 0, c = dx, d = dy;
        if (in.ipotcn == 1) { a = in.xl + (ix0 - 1) * dx; b = in.yl + (iy0 - 1) * dy; }
        if (in.ipotcn == 2) {
            a = in.xl + (ix0 - 1) * dx; b = in.yl + (iy0 - 1) * dy;
            c = in.xh - (mx - ix1) * dx; d = in.yh - (my - iy1) * dy;
        }
        if (in.ipotcn == 3) { a = in.xc1 + (ix0 - 1) * dx; b = in.yc1 + (iy0 - 1) * dy; }
        if (in.ipotcn == 4) {
            a = in.xc1 + (ix0 - 1) * dx; b = in.yc1 + (iy0 - 1) * dy;
            c = in.xcm - (mx - ix1) * dx; d = in.ycm - (my - iy1) * dy;
        }
        std::fprintf(f, "%s%s%s%s%s%s%sMX,MY,XL,YL,DX,DY\n", fort_i(ix1 - ix0 + 1, 5).c_str(),
                     fort_i(iy1 - iy0 + 1, 5).c_str(), fort_g(a, 12, 4).c_str(), fort_g(b, 12, 4).c_str(),
                     fort_g(c, 12, 4).c_str(), fort_g(d, 12, 4).c_str(), "");
    }
    }

    // undeformed distance (Z = 2, grid forms)
    if (in.ipotcn > 0) {
        std::fprintf(f, "%s%s%48sIBASE, IPLAN\n", fort_i(in.ibase, 5).c_str(), fort_i(1, 5).c_str(), "");
        if (in.ibase == 1) {
            std::fprintf(f, "%%  UNDEF.DISTANCE, (1)-(2): QUADRATIC\n");
            std::string l1;
            for (int k = 0; k < 5; k++) l1 += fort_g(in.prmudf[k], 12, 4);
            std::fprintf(f, "%s\n%s%46sB(k), k=1, 6\n", l1.c_str(), fort_g(in.prmudf[5], 12, 4).c_str(), "");
        } else if (in.ibase == 9) {
            std::fprintf(f, "%%  UNDEF.DISTANCE, (1)-(2): SPECIFIED PER ELEMENT\n");
            std::vector<double> row;
            for (int iy = iy0; iy <= iy1; iy++) {
                row.clear();
                for (int ix = ix0; ix <= ix1; ix++) {
                    const size_t ii = (size_t)(iy - 1) * mx + (ix - 1);
                    row.push_back(ii < in.h_elem.size() ? in.h_elem[ii] : 0.0);
                }
                if (iy > 1) std::fprintf(f, " \n");
                write_row_7509(f, row.data(), (int)row.size());
            }
        }
        std::fprintf(f, "%%  UNRESTRICTED PLANFORM\n");
    }
    std::fprintf(f, " \n");
}

}  // namespace cntc
