// bench_mass: the reference's mass benchmark driver interface on the GPU.
#include "driver_common.hpp"

int main(int argc, char** argv) { return driver::main_impl(false, argc, argv); }
