// fp32_shifted.cpp — FP32 shifted-storage mode (populations stored as FP32
// deviations from a rest equilibrium).  Placeholder entry points until the
// FP32 kernels land: every call reports DRP_ERR_UNSUPPORTED.
#include "lattice_impl.h"

static int unsupported() { return drp_fail(DRP_ERR_UNSUPPORTED, "DRP_FP32_SHIFTED is not implemented yet"); }

int fp32_init_from_host_fields(drp_lattice*, const double* const*) { return unsupported(); }
int fp32_step(drp_lattice*, int) { return unsupported(); }
int fp32_get_macro(drp_lattice*, int, double*) { return unsupported(); }
int fp32_get_populations(drp_lattice*, int, double*) { return unsupported(); }
int fp32_set_populations(drp_lattice*, int, const double*, int) { return unsupported(); }
int fp32_macro_fp64_view(drp_lattice*, const double**) { return unsupported(); }
