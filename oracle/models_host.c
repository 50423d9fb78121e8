/*
 * models_host.c - TEST INFRASTRUCTURE ONLY.
 *
 * Builds the agent models of tay_amd/agent_src as host C callbacks for the
 * oracle, the way the reference builds its models for the CPU path
 * (test/agent_host.h:1-27, test/agent_host.c): the OpenCL address-space
 * qualifiers are defined away and the packed layouts come from tay.h.
 */
#include "tay.h"
#include <math.h>

#define global
#define constant

#include "taystd.h"
#include "test_agent.h"
#include "flocking.h"
#include "sph.h"

#include "taystd.c"
#include "test_agent.c"
#include "flocking.c"
#include "sph.c"
