// src/contact.h of the reference: the penalty contact the reference's main
// loop applies between bodies after each step (src/mfree_iwf.cpp:91-97).
#ifndef MFREE_CONTACT_H_
#define MFREE_CONTACT_H_

#include "body.h"

template <typename T, typename Q>
void contact_apply_body_to_boundary(body<T>& master, body<Q>& slave) {
  if (!(slave.has_boundary() && master.has_boundary())) return;
  mfree_adapter::die("contact_apply_body_to_boundary: penalty contact between bodies is not part of this library");
}

#endif
