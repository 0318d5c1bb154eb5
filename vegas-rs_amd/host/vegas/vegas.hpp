// Host library of the MI355X Metropolis engine: the reference's Rust trait
// surface (state.rs, thermostat.rs, energy.rs, integrator.rs, machine.rs,
// instrument.rs, accumulator.rs, program.rs) in C++ over the C ABI of
// libvegas_hip.so, and a job server for many programs at once.
#pragma once

#include "vegas_hip.h"

#include "accumulator.hpp"
#include "checkpoint.hpp"
#include "energy.hpp"
#include "error.hpp"
#include "instrument.hpp"
#include "lattice.hpp"
#include "machine.hpp"
#include "matrix.hpp"
#include "parquet.hpp"
#include "program.hpp"
#include "server.hpp"
#include "snappy.hpp"
#include "state.hpp"
#include "thermostat.hpp"
