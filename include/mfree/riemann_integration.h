// src/riemann_integration.h of the reference: its actions are declared, with every other
// action of the reference, in actions.h (through body.h) and run inside
// body<T>::step() on the GPU.
#ifndef MFREE_RIEMANN_INTEGRATION_H_
#define MFREE_RIEMANN_INTEGRATION_H_

#include "body.h"

#endif
