// inp_writer.hpp -- CONTACT's input-file description of a module-3 case (the .inp file that
// the flag CNTC_if_wrtinp asks for, so that a case solved through the library can be replayed
// by the stand-alone program).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "engine.hpp"

namespace cntc {

// Fortran output edit descriptors as the reference's formats use them: Iw.m, Fw.d, Gw.d
// (F or E form by magnitude, Fortran 2008 10.7.5.2.2), Ew.d and ESw.d.
std::string fort_i(long v, int w, int m = 0);
std::string fort_f(double x, int w, int d);
std::string fort_g(double x, int w, int d);
std::string fort_e(double x, int w, int d);
std::string fort_es(double x, int w, int d);

// Control digits of a case that the library keeps without solving with them (stored through
// cntc_setflags), with the defaults of src/m_hierarch_data.f90:970-1010.
struct InpDigits {
    int stress = 0, frclaw = 0, mater = 0, matfil = 1, output = 2, flow = 4, ret = 1;
};

// Writes case `ncase` of problem (ire, icp) in the form of wrtinp (src/m_sinput.f90:1506-1862)
// after the header line of cntc_calculate (src/contact_addon.f90:3860-3867).  `el` is the
// element division of the solution (mx*my) or empty when there is none; it sets the box
// around the contact area to which grid forms and element-wise gaps are clipped.
void write_inp_case(std::FILE* f, int ire, int icp, int ncase, const Patch& P, const InpDigits& dg,
                    double fkin, const std::vector<int8_t>& el);

}  // namespace cntc
