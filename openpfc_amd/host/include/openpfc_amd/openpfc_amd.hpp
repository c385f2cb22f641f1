// openpfc_amd.hpp -- umbrella header of the C++ host layer.
#pragma once

#include "autotune.hpp"
#include "comm.hpp"
#include "core.hpp"
#include "decomposition.hpp"
#include "fft.hpp"
#include "field_modifier.hpp"
#include "io.hpp"
#include "model.hpp"
#include "pfc_model_hip.hpp"
#include "tungsten_model_hip.hpp"
#include "simulator.hpp"
#include "time.hpp"
#include "world.hpp"
