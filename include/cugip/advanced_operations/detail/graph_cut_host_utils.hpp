// Same path as the reference's detail/graph_cut_host_utils.hpp (GraphData<T>); GraphData is
// defined in cugip/graph_cut.hpp.
#pragma once

#include "../../graph_cut.hpp"
