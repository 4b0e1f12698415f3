// qsv_fused.hip -- placeholder: dense blocks on C64 states use the FP64 block kernel.
#include "qsv_internal.h"

namespace qsv {

int launch_block_mfma(qsv_state *s, const int *qubits, int k, const double *u, int)
{
    return launch_block(s, qubits, k, u, 0);
}

}  // namespace qsv
