// Reader for flat structural Verilog netlists over the AIGPDK cell library
// (aigpdk/aigpdk.v): INV, BUF, AND2_{00,01,10,11}_0, AND2_11_1, DFF, DFFSR;
// GEM_ASSERT / GEM_DISPLAY instances are accepted and ignored.  Builds the
// AIG with the reference's cell semantics (src/aig.rs:1044-1064, 2028-2055):
// every input port becomes an input pin, every clock port driving a DFF CLK
// also gets a posedge flag pin, a DFF's enable is that flag.  Port metadata
// for VCD I/O is stored in the Aig's port view.
#pragma once

#include <string>

#include "aig.hpp"

namespace jq {

void read_aigpdk_verilog(const std::string &text, Aig &g);

}  // namespace jq
