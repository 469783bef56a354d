/*
 * Izhikevich2007DopSyn plugin (libIzhikevich2007DopSyn.so), bit-exact with src/nemo/cpu/plugins/Izhikevich2007DopSyn.cpp:14-156
 * (the update is in variants.h).
 */
#define NEMO_VARIANT VARIANT_DOPSYN
#include "variants.h"
