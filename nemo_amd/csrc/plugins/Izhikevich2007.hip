/*
 * Izhikevich2007 plugin (libIzhikevich2007.so), bit-exact with src/nemo/cpu/plugins/Izhikevich2007.cpp:14-134
 * (the update is in variants.h).
 */
#define NEMO_VARIANT VARIANT_2007
#include "variants.h"
