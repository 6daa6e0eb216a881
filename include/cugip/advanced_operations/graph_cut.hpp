// Same path as the reference's include/cugip/advanced_operations/graph_cut.hpp, so its callers
// (test/graph_cut.cu, benchmarks/graph_cut/graph_cut.cu) switch by changing the include directory.
#pragma once

#include "../graph_cut.hpp"
