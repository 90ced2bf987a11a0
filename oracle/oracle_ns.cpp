// placeholder, replaced below
#include "oracle.h"
extern "C" {
int orc_ns_step(const orc_mesh*, const orc_ns_params*, double, const double*, double*, double*, double*, double*) { return -1; }
double orc_ns_dt(const orc_mesh*, const orc_ns_params*, const double*, double) { return 0.0; }
}
