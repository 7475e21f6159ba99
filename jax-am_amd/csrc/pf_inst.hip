// Register-tiled fused-step instantiations.  Compiled once per part with
//   -DPF_INST_DTYPE=<PF_F64|PF_F32|PF_BF16 value> -DPF_KLO=<k> -DPF_KHI=<k> -DPF_PART=<n>
// (see Makefile) so the instantiations build in parallel.  K = 1..24 and 32 have
// register-tiled kernels; other K run on the generic kernel.
// Tile shapes (TX, TY): F32/BF16: 0 = 32x8, 1 = 64x4, 2 = 16x8;  F64: 0 = 32x4, 1 = 16x4.
#include "pf_kernels.h"

#if !defined(PF_INST_DTYPE) || !defined(PF_KLO) || !defined(PF_KHI) || !defined(PF_PART)
#error "define PF_INST_DTYPE, PF_KLO, PF_KHI and PF_PART"
#endif

#define PF_CAT2(a, b) a##b
#define PF_CAT(a, b) PF_CAT2(a, b)

#if PF_INST_DTYPE == 1
#define PF_K_ENTRIES(K)                                                                          \
  PF_ENTRY(PF_F32, float, K, 0, 0, 32, 8), PF_ENTRY(PF_F32, float, K, 0, 1, 64, 4),             \
      PF_ENTRY(PF_F32, float, K, 0, 2, 16, 8), PF_ENTRY(PF_F32, float, K, 1, 0, 32, 8),         \
      PF_ENTRY(PF_F32, float, K, 1, 1, 64, 4), PF_ENTRY(PF_F32, float, K, 1, 2, 16, 8)
#elif PF_INST_DTYPE == 2
#define PF_K_ENTRIES(K)                                                                          \
  PF_ENTRY(PF_BF16, bf16_t, K, 0, 0, 32, 8), PF_ENTRY(PF_BF16, bf16_t, K, 0, 1, 64, 4),         \
      PF_ENTRY(PF_BF16, bf16_t, K, 0, 2, 16, 8), PF_ENTRY(PF_BF16, bf16_t, K, 1, 0, 32, 8),     \
      PF_ENTRY(PF_BF16, bf16_t, K, 1, 1, 64, 4), PF_ENTRY(PF_BF16, bf16_t, K, 1, 2, 16, 8)
#elif PF_INST_DTYPE == 0
#define PF_K_ENTRIES(K)                                                                          \
  PF_ENTRY(PF_F64, double, K, 0, 0, 32, 4), PF_ENTRY(PF_F64, double, K, 0, 1, 16, 4),           \
      PF_ENTRY(PF_F64, double, K, 1, 0, 32, 4), PF_ENTRY(PF_F64, double, K, 1, 1, 16, 4)
#else
#error "unknown PF_INST_DTYPE"
#endif

namespace pf {

static const StepEntry kTable[] = {
#if PF_KLO <= 1 && 1 <= PF_KHI
    PF_K_ENTRIES(1),
#endif
#if PF_KLO <= 2 && 2 <= PF_KHI
    PF_K_ENTRIES(2),
#endif
#if PF_KLO <= 3 && 3 <= PF_KHI
    PF_K_ENTRIES(3),
#endif
#if PF_KLO <= 4 && 4 <= PF_KHI
    PF_K_ENTRIES(4),
#endif
#if PF_KLO <= 5 && 5 <= PF_KHI
    PF_K_ENTRIES(5),
#endif
#if PF_KLO <= 6 && 6 <= PF_KHI
    PF_K_ENTRIES(6),
#endif
#if PF_KLO <= 7 && 7 <= PF_KHI
    PF_K_ENTRIES(7),
#endif
#if PF_KLO <= 8 && 8 <= PF_KHI
    PF_K_ENTRIES(8),
#endif
#if PF_KLO <= 9 && 9 <= PF_KHI
    PF_K_ENTRIES(9),
#endif
#if PF_KLO <= 10 && 10 <= PF_KHI
    PF_K_ENTRIES(10),
#endif
#if PF_KLO <= 11 && 11 <= PF_KHI
    PF_K_ENTRIES(11),
#endif
#if PF_KLO <= 12 && 12 <= PF_KHI
    PF_K_ENTRIES(12),
#endif
#if PF_KLO <= 13 && 13 <= PF_KHI
    PF_K_ENTRIES(13),
#endif
#if PF_KLO <= 14 && 14 <= PF_KHI
    PF_K_ENTRIES(14),
#endif
#if PF_KLO <= 15 && 15 <= PF_KHI
    PF_K_ENTRIES(15),
#endif
#if PF_KLO <= 16 && 16 <= PF_KHI
    PF_K_ENTRIES(16),
#endif
#if PF_KLO <= 17 && 17 <= PF_KHI
    PF_K_ENTRIES(17),
#endif
#if PF_KLO <= 18 && 18 <= PF_KHI
    PF_K_ENTRIES(18),
#endif
#if PF_KLO <= 19 && 19 <= PF_KHI
    PF_K_ENTRIES(19),
#endif
#if PF_KLO <= 20 && 20 <= PF_KHI
    PF_K_ENTRIES(20),
#endif
#if PF_KLO <= 21 && 21 <= PF_KHI
    PF_K_ENTRIES(21),
#endif
#if PF_KLO <= 22 && 22 <= PF_KHI
    PF_K_ENTRIES(22),
#endif
#if PF_KLO <= 23 && 23 <= PF_KHI
    PF_K_ENTRIES(23),
#endif
#if PF_KLO <= 24 && 24 <= PF_KHI
    PF_K_ENTRIES(24),
#endif
#if PF_KLO <= 32 && 32 <= PF_KHI
    PF_K_ENTRIES(32),
#endif
};

const StepEntry* PF_CAT(step_table_part_, PF_PART)(int* n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace pf
