// src/cont_mech.h of the reference: its actions are declared, with every other
// action of the reference, in actions.h (through body.h) and run inside
// body<T>::step() on the GPU.
#ifndef MFREE_CONT_MECH_H_
#define MFREE_CONT_MECH_H_

#include "body.h"

#endif
