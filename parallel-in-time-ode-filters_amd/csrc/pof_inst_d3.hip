// Kernel instantiations for ODE dimension d = 3.
#include "pof_kernels.hpp"

namespace pof {

const EngineFns* engines_d3(int D) {
  static const EngineFns table[] = {
      EngineFns{6, 3, &run_iteration<6, 3>, &run_smooth_chol<6, 3>,
                &run_output<6, 3>, &run_filtered<6, 3>,
                &run_seq_iteration<6, 3>, &run_seq_smooth_chol<6, 3>, &run_smooth_only<6, 3>,
                &run_linearize<6, 3>,
                &run_linearize_alt<6, 3>},
      EngineFns{9, 3, &run_iteration<9, 3>, &run_smooth_chol<9, 3>,
                &run_output<9, 3>, &run_filtered<9, 3>,
                &run_seq_iteration<9, 3>, &run_seq_smooth_chol<9, 3>, &run_smooth_only<9, 3>,
                &run_linearize<9, 3>,
                &run_linearize_alt<9, 3>},
      EngineFns{12, 3, &run_iteration<12, 3>, &run_smooth_chol<12, 3>,
                &run_output<12, 3>, &run_filtered<12, 3>,
                &run_seq_iteration<12, 3>, &run_seq_smooth_chol<12, 3>, &run_smooth_only<12, 3>,
                &run_linearize<12, 3>,
                &run_linearize_alt<12, 3>},
      EngineFns{15, 3, &run_iteration<15, 3>, &run_smooth_chol<15, 3>,
                &run_output<15, 3>, &run_filtered<15, 3>,
                &run_seq_iteration<15, 3>, &run_seq_smooth_chol<15, 3>, &run_smooth_only<15, 3>,
                &run_linearize<15, 3>,
                &run_linearize_alt<15, 3>},
      EngineFns{18, 3, &run_iteration<18, 3>, &run_smooth_chol<18, 3>,
                &run_output<18, 3>, &run_filtered<18, 3>,
                &run_seq_iteration<18, 3>, &run_seq_smooth_chol<18, 3>, &run_smooth_only<18, 3>,
                &run_linearize<18, 3>,
                &run_linearize_alt<18, 3>},
  };
  for (const auto& e : table)
    if (e.D == D) return &e;
  return nullptr;
}

size_t lds_bytes_d3(int D) {
  switch (D) {
    case 6: return Geo<6, kLdsFull>::lds_bytes();
    case 9: return Geo<9, kLdsFull>::lds_bytes();
    case 12: return Geo<12, kLdsFull>::lds_bytes();
    case 15: return Geo<15, kLdsFull>::lds_bytes();
    case 18: return Geo<18, kLdsFull>::lds_bytes();
    default: return 0;
  }
}

}  // namespace pof
