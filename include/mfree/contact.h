// src/contact.h of the reference: the penalty contact the reference's main
// loop applies between bodies after each step (src/mfree_iwf.cpp:91-97), run
// on the GPU (mfree_contact): the slave body's contact forces, which its
// next step reads.
#ifndef MFREE_CONTACT_H_
#define MFREE_CONTACT_H_

#include "body.h"

template <typename T, typename Q>
void contact_apply_body_to_boundary(body<T>& master, body<Q>& slave) {
  if (!(slave.has_boundary() && master.has_boundary())) return;
  if (!master.context() || !slave.context())
    mfree_adapter::die("contact_apply_body_to_boundary: both bodies must have stepped (their state is on the GPU)");
  if (mfree_contact(master.context(), 0, slave.context(), 0) != MFREE_OK)
    mfree_adapter::die(std::string("contact_apply_body_to_boundary: ") + mfree_last_error(slave.context()));
  slave.mark_stale();
}

#endif
