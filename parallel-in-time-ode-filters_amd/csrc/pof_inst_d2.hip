// Kernel instantiations for ODE dimension d = 2.
#include "pof_kernels.hpp"

namespace pof {

const EngineFns* engines_d2(int D) {
  static const EngineFns table[] = {
      EngineFns{4, 2, &run_iteration<4, 2>, &run_smooth_chol<4, 2>,
                &run_output<4, 2>, &run_filtered<4, 2>,
                &run_seq_iteration<4, 2>, &run_seq_smooth_chol<4, 2>, &run_smooth_only<4, 2>,
                &run_linearize<4, 2>,
                &run_linearize_alt<4, 2>},
      EngineFns{6, 2, &run_iteration<6, 2>, &run_smooth_chol<6, 2>,
                &run_output<6, 2>, &run_filtered<6, 2>,
                &run_seq_iteration<6, 2>, &run_seq_smooth_chol<6, 2>, &run_smooth_only<6, 2>,
                &run_linearize<6, 2>,
                &run_linearize_alt<6, 2>},
      EngineFns{8, 2, &run_iteration<8, 2>, &run_smooth_chol<8, 2>,
                &run_output<8, 2>, &run_filtered<8, 2>,
                &run_seq_iteration<8, 2>, &run_seq_smooth_chol<8, 2>, &run_smooth_only<8, 2>,
                &run_linearize<8, 2>,
                &run_linearize_alt<8, 2>},
      EngineFns{10, 2, &run_iteration<10, 2>, &run_smooth_chol<10, 2>,
                &run_output<10, 2>, &run_filtered<10, 2>,
                &run_seq_iteration<10, 2>, &run_seq_smooth_chol<10, 2>, &run_smooth_only<10, 2>,
                &run_linearize<10, 2>,
                &run_linearize_alt<10, 2>},
      EngineFns{12, 2, &run_iteration<12, 2>, &run_smooth_chol<12, 2>,
                &run_output<12, 2>, &run_filtered<12, 2>,
                &run_seq_iteration<12, 2>, &run_seq_smooth_chol<12, 2>, &run_smooth_only<12, 2>,
                &run_linearize<12, 2>,
                &run_linearize_alt<12, 2>},
  };
  for (const auto& e : table)
    if (e.D == D) return &e;
  return nullptr;
}

size_t lds_bytes_d2(int D) {
  switch (D) {
    case 4: return Geo<4, kLdsFull>::lds_bytes();
    case 6: return Geo<6, kLdsFull>::lds_bytes();
    case 8: return Geo<8, kLdsFull>::lds_bytes();
    case 10: return Geo<10, kLdsFull>::lds_bytes();
    case 12: return Geo<12, kLdsFull>::lds_bytes();
    default: return 0;
  }
}

}  // namespace pof
