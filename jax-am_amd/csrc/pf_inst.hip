// Register-tiled fused-step instantiations.  Compiled once per part with
//   -DPF_INST_DTYPE=<PF_F64|PF_F32|PF_BF16 value> -DPF_KLO=<k> -DPF_KHI=<k> -DPF_PART=<n>
// (see Makefile) so the instantiations build in parallel.  K = 1..24 and 32 have
// register-tiled kernels; other K run on the generic kernel.
// Tile shapes (TX x TY cells, H threads per cell):
//   F32/BF16: 0 = 32x8, 1 = 64x4, 2 = 16x8; even K also 3 = 16x8 H2, 4 = 8x8 H2
//   F64:      0 = 32x4, 1 = 16x4;           even K also 2 = 16x4 H2, 3 = 8x4 H2
#include "pf_kernels.h"

#if !defined(PF_INST_DTYPE) || !defined(PF_KLO) || !defined(PF_KHI) || !defined(PF_PART)
#error "define PF_INST_DTYPE, PF_KLO, PF_KHI and PF_PART"
#endif

#define PF_CAT2(a, b) a##b
#define PF_CAT(a, b) PF_CAT2(a, b)

#if PF_INST_DTYPE == 1 || PF_INST_DTYPE == 2
#if PF_INST_DTYPE == 1
#define PF_DT PF_F32
#define PF_TS float
#else
#define PF_DT PF_BF16
#define PF_TS bf16_t
#endif
#define PF_K_ODD(K)                                                                              \
  PF_ENTRY(PF_DT, PF_TS, K, 0, 0, 32, 8, 1), PF_ENTRY(PF_DT, PF_TS, K, 0, 1, 64, 4, 1),           \
      PF_ENTRY(PF_DT, PF_TS, K, 0, 2, 16, 8, 1), PF_ENTRY(PF_DT, PF_TS, K, 1, 0, 32, 8, 1),       \
      PF_ENTRY(PF_DT, PF_TS, K, 1, 1, 64, 4, 1), PF_ENTRY(PF_DT, PF_TS, K, 1, 2, 16, 8, 1)
#define PF_K_EVEN(K)                                                                             \
  PF_K_ODD(K), PF_ENTRY(PF_DT, PF_TS, K, 0, 3, 16, 8, 2), PF_ENTRY(PF_DT, PF_TS, K, 0, 4, 8, 8, 2), \
      PF_ENTRY(PF_DT, PF_TS, K, 1, 3, 16, 8, 2), PF_ENTRY(PF_DT, PF_TS, K, 1, 4, 8, 8, 2)
#elif PF_INST_DTYPE == 0
#define PF_K_ODD(K)                                                                              \
  PF_ENTRY(PF_F64, double, K, 0, 0, 32, 4, 1), PF_ENTRY(PF_F64, double, K, 0, 1, 16, 4, 1),       \
      PF_ENTRY(PF_F64, double, K, 1, 0, 32, 4, 1), PF_ENTRY(PF_F64, double, K, 1, 1, 16, 4, 1)
#define PF_K_EVEN(K)                                                                             \
  PF_K_ODD(K), PF_ENTRY(PF_F64, double, K, 0, 2, 16, 4, 2), PF_ENTRY(PF_F64, double, K, 0, 3, 8, 4, 2), \
      PF_ENTRY(PF_F64, double, K, 1, 2, 16, 4, 2), PF_ENTRY(PF_F64, double, K, 1, 3, 8, 4, 2)
#else
#error "unknown PF_INST_DTYPE"
#endif

namespace pf {

static const StepEntry kTable[] = {
#if PF_KLO <= 1 && 1 <= PF_KHI
    PF_K_ODD(1),
#endif
#if PF_KLO <= 2 && 2 <= PF_KHI
    PF_K_EVEN(2),
#endif
#if PF_KLO <= 3 && 3 <= PF_KHI
    PF_K_ODD(3),
#endif
#if PF_KLO <= 4 && 4 <= PF_KHI
    PF_K_EVEN(4),
#endif
#if PF_KLO <= 5 && 5 <= PF_KHI
    PF_K_ODD(5),
#endif
#if PF_KLO <= 6 && 6 <= PF_KHI
    PF_K_EVEN(6),
#endif
#if PF_KLO <= 7 && 7 <= PF_KHI
    PF_K_ODD(7),
#endif
#if PF_KLO <= 8 && 8 <= PF_KHI
    PF_K_EVEN(8),
#endif
#if PF_KLO <= 9 && 9 <= PF_KHI
    PF_K_ODD(9),
#endif
#if PF_KLO <= 10 && 10 <= PF_KHI
    PF_K_EVEN(10),
#endif
#if PF_KLO <= 11 && 11 <= PF_KHI
    PF_K_ODD(11),
#endif
#if PF_KLO <= 12 && 12 <= PF_KHI
    PF_K_EVEN(12),
#endif
#if PF_KLO <= 13 && 13 <= PF_KHI
    PF_K_ODD(13),
#endif
#if PF_KLO <= 14 && 14 <= PF_KHI
    PF_K_EVEN(14),
#endif
#if PF_KLO <= 15 && 15 <= PF_KHI
    PF_K_ODD(15),
#endif
#if PF_KLO <= 16 && 16 <= PF_KHI
    PF_K_EVEN(16),
#endif
#if PF_KLO <= 17 && 17 <= PF_KHI
    PF_K_ODD(17),
#endif
#if PF_KLO <= 18 && 18 <= PF_KHI
    PF_K_EVEN(18),
#endif
#if PF_KLO <= 19 && 19 <= PF_KHI
    PF_K_ODD(19),
#endif
#if PF_KLO <= 20 && 20 <= PF_KHI
    PF_K_EVEN(20),
#endif
#if PF_KLO <= 21 && 21 <= PF_KHI
    PF_K_ODD(21),
#endif
#if PF_KLO <= 22 && 22 <= PF_KHI
    PF_K_EVEN(22),
#endif
#if PF_KLO <= 23 && 23 <= PF_KHI
    PF_K_ODD(23),
#endif
#if PF_KLO <= 24 && 24 <= PF_KHI
    PF_K_EVEN(24),
#endif
#if PF_KLO <= 32 && 32 <= PF_KHI
    PF_K_EVEN(32),
#endif
};

const StepEntry* PF_CAT(step_table_part_, PF_PART)(int* n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace pf
