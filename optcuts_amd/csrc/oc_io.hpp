// Trace lines of the Newton loop in OptCuts' energyValPerIter.txt and
// gradientPerIter.txt formats (host code, oc_io.cpp).
#pragma once

#include <cstdio>

namespace oc {
namespace io {

// "E_total E_SD seam a0" (Optimizer.cpp:706-728 with one energy term)
int trace_energy_line(FILE* fp, double Etotal, double Esd, double seam, double a0);
// "|g|^2 |g_SD|^2" (Optimizer.cpp:729-739 with one energy term)
int trace_grad_line(FILE* fp, double gsq, double gsqSD);

}  // namespace io
}  // namespace oc
