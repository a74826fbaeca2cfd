// lambda_q8 for builds without AVM.
//
// CelV2 computes lambda_q8(qindex) = llround(0.002 * q^2 * (q120 / q)^0.8 * 256) with q AVM's AC
// quantiser step av2_ac_quant_QTX(qindex, 0, 0, AVM_BITS_8) / 8 (src/core/quant.cpp:31-39). AVM's
// table is not part of this tree, so celv2_lambda_q8 evaluates the same formula on a stated model of
// the step, q(qindex) = (4 + qindex + qindex^3 / 9000) / 8 (monotone, 0.5 at 0, about 263 at 255).
// The engines take lambda as an input; the HIP and CPU engines agree for any table.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
int64_t celv2_lambda_q8(int qindex);
#ifdef __cplusplus
}
#endif
