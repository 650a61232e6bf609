// koo_host.h -- declarations shared by the host-only translation units.
#pragma once

namespace koo {
// Records a message for koo_last_error() and returns the status.
int set_error(int status, const char* fmt, ...);
void exp_corr(const double* times, int n, double L, double beta, double* corr);
void pseudo_sqrt(int n, const double* cov, int factors, bool normalise, double* out);
}  // namespace koo

#define set_host_error koo::set_error
