/*
 * IzhikevichChaos plugin (libIzhikevichChaos.so), bit-exact with src/nemo/cpu/plugins/IzhikevichChaos.cpp:14-101
 * (the update is in variants.h).
 */
#define NEMO_VARIANT VARIANT_CHAOS
#include "variants.h"
