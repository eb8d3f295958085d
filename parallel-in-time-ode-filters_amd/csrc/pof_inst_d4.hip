// Kernel instantiations for ODE dimension d = 4.
#include "pof_kernels.hpp"

namespace pof {

const EngineFns* engines_d4(int D) {
  static const EngineFns table[] = {
      EngineFns{8, 4, &run_iteration<8, 4>, &run_smooth_chol<8, 4>,
                &run_output<8, 4>, &run_filtered<8, 4>,
                &run_seq_iteration<8, 4>, &run_seq_smooth_chol<8, 4>, &run_smooth_only<8, 4>,
                &run_linearize<8, 4>,
                &run_linearize_alt<8, 4>},
      EngineFns{12, 4, &run_iteration<12, 4>, &run_smooth_chol<12, 4>,
                &run_output<12, 4>, &run_filtered<12, 4>,
                &run_seq_iteration<12, 4>, &run_seq_smooth_chol<12, 4>, &run_smooth_only<12, 4>,
                &run_linearize<12, 4>,
                &run_linearize_alt<12, 4>},
      EngineFns{16, 4, &run_iteration<16, 4>, &run_smooth_chol<16, 4>,
                &run_output<16, 4>, &run_filtered<16, 4>,
                &run_seq_iteration<16, 4>, &run_seq_smooth_chol<16, 4>, &run_smooth_only<16, 4>,
                &run_linearize<16, 4>,
                &run_linearize_alt<16, 4>},
      EngineFns{20, 4, &run_iteration<20, 4>, &run_smooth_chol<20, 4>,
                &run_output<20, 4>, &run_filtered<20, 4>,
                &run_seq_iteration<20, 4>, &run_seq_smooth_chol<20, 4>, &run_smooth_only<20, 4>,
                &run_linearize<20, 4>,
                &run_linearize_alt<20, 4>},
  };
  for (const auto& e : table)
    if (e.D == D) return &e;
  return nullptr;
}

size_t lds_bytes_d4(int D) {
  switch (D) {
    case 8: return Geo<8, kLdsFull>::lds_bytes();
    case 12: return Geo<12, kLdsFull>::lds_bytes();
    case 16: return Geo<16, kLdsFull>::lds_bytes();
    case 20: return Geo<20, kLdsFull>::lds_bytes();
    default: return 0;
  }
}

}  // namespace pof
