// oracle_flops.cpp -- the oracle (oracle.cpp) compiled with a counting float64
// type, to measure the algorithmic FP64 work per neutron of a chain.
// MEASUREMENT INFRASTRUCTURE: feeds the roofline figure of bench.py.
//
// Every system header is included first; then `double` is redefined to the
// counting type for the oracle's own code only.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

#include "../include/mcv_acc.h"
#include "flopcount.h"

namespace orcflops {
thread_local Counters g;
}

#define ORC_FLOP_BUILD 1
#define double orcflops::CReal
#include "oracle.cpp"
#undef double

extern "C" {

void orc_flops_reset(void) {
  orcflops::g.arith = 0;
  orcflops::g.trans = 0;
  orcflops::g.draws = 0;
  orcflops::g.paused = 0;
}

// out[3] = arithmetic ops, transcendental calls, random draws (this thread)
void orc_flops_read(uint64_t* out3) {
  out3[0] = orcflops::g.arith;
  out3[1] = orcflops::g.trans;
  out3[2] = orcflops::g.draws;
}

}  // extern "C"
