/*
 * IzhikevichSyn plugin (libIzhikevichSyn.so), bit-exact with src/nemo/cpu/plugins/IzhikevichSyn.cpp:14-140
 * (the update is in variants.h).
 */
#define NEMO_VARIANT VARIANT_SYN
#include "variants.h"
