// Kernel instantiations for ODE dimension d = 1.
#include "pof_kernels.hpp"

namespace pof {

const EngineFns* engines_d1(int D) {
  static const EngineFns table[] = {
      EngineFns{1, 1, &run_iteration<1, 1>, &run_smooth_chol<1, 1>,
                &run_output<1, 1>, &run_filtered<1, 1>,
                &run_seq_iteration<1, 1>, &run_seq_smooth_chol<1, 1>, &run_smooth_only<1, 1>,
                &run_linearize<1, 1>,
                &run_linearize_alt<1, 1>},
      EngineFns{2, 1, &run_iteration<2, 1>, &run_smooth_chol<2, 1>,
                &run_output<2, 1>, &run_filtered<2, 1>,
                &run_seq_iteration<2, 1>, &run_seq_smooth_chol<2, 1>, &run_smooth_only<2, 1>,
                &run_linearize<2, 1>,
                &run_linearize_alt<2, 1>},
      EngineFns{3, 1, &run_iteration<3, 1>, &run_smooth_chol<3, 1>,
                &run_output<3, 1>, &run_filtered<3, 1>,
                &run_seq_iteration<3, 1>, &run_seq_smooth_chol<3, 1>, &run_smooth_only<3, 1>,
                &run_linearize<3, 1>,
                &run_linearize_alt<3, 1>},
      EngineFns{4, 1, &run_iteration<4, 1>, &run_smooth_chol<4, 1>,
                &run_output<4, 1>, &run_filtered<4, 1>,
                &run_seq_iteration<4, 1>, &run_seq_smooth_chol<4, 1>, &run_smooth_only<4, 1>,
                &run_linearize<4, 1>,
                &run_linearize_alt<4, 1>},
      EngineFns{5, 1, &run_iteration<5, 1>, &run_smooth_chol<5, 1>,
                &run_output<5, 1>, &run_filtered<5, 1>,
                &run_seq_iteration<5, 1>, &run_seq_smooth_chol<5, 1>, &run_smooth_only<5, 1>,
                &run_linearize<5, 1>,
                &run_linearize_alt<5, 1>},
      EngineFns{6, 1, &run_iteration<6, 1>, &run_smooth_chol<6, 1>,
                &run_output<6, 1>, &run_filtered<6, 1>,
                &run_seq_iteration<6, 1>, &run_seq_smooth_chol<6, 1>, &run_smooth_only<6, 1>,
                &run_linearize<6, 1>,
                &run_linearize_alt<6, 1>},
  };
  for (const auto& e : table)
    if (e.D == D) return &e;
  return nullptr;
}

size_t lds_bytes_d1(int D) {
  switch (D) {
    case 1: return Geo<1, kLdsFull>::lds_bytes();
    case 2: return Geo<2, kLdsFull>::lds_bytes();
    case 3: return Geo<3, kLdsFull>::lds_bytes();
    case 4: return Geo<4, kLdsFull>::lds_bytes();
    case 5: return Geo<5, kLdsFull>::lds_bytes();
    case 6: return Geo<6, kLdsFull>::lds_bytes();
    default: return 0;
  }
}

}  // namespace pof
