/*
 * oracle_filter.c -- TEST INFRASTRUCTURE ONLY (the CPU checker).
 * Restatement of the multi-scale filter sensor, Filter.cpp:12-1720.
 */
#include "oracle_common.h"
