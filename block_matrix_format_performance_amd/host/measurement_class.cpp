// measurement_class.cpp -- measurement_class constructor
// (reference: common/measurement_class.cpp:7-16).
#include "bmfp/measurement_class.h"

measurement_class::measurement_class(std::string format, double elapsed_arg, double load_store_bytes,
                                     double operations_count)
  : elapsed(elapsed_arg)
  , effective_bandwidth(load_store_bytes / (elapsed_arg * giga))
  , computational_throughput(operations_count / (elapsed_arg * giga))
  , matrix_format(std::move(format))
{
}
