// Registry of the kernel instantiations (one translation unit per ODE dimension).
#include "pof_engine.hpp"

namespace pof {

const EngineFns* engines_d1(int D);
const EngineFns* engines_d2(int D);
const EngineFns* engines_d3(int D);
const EngineFns* engines_d4(int D);
size_t lds_bytes_d1(int D);
size_t lds_bytes_d2(int D);
size_t lds_bytes_d3(int D);
size_t lds_bytes_d4(int D);

const EngineFns* find_engine(int D, int d) {
  switch (d) {
    case 1: return engines_d1(D);
    case 2: return engines_d2(D);
    case 3: return engines_d3(D);
    case 4: return engines_d4(D);
    default: return nullptr;
  }
}

}  // namespace pof
