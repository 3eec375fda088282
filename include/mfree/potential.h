// src/potential.h of the reference: its actions are declared, with every other
// action of the reference, in actions.h (through body.h) and run inside
// body<T>::step() on the GPU.
#ifndef MFREE_POTENTIAL_H_
#define MFREE_POTENTIAL_H_

#include "body.h"

#endif
