// bench_stiffness: the reference's stiffness benchmark driver interface on the
// GPU.  The table holds [phi | dphi] (2 (P+1)^2 values).
#include "driver_common.hpp"

int main(int argc, char** argv) { return driver::main_impl(true, argc, argv); }
