#pragma once

namespace tsg {

// sign of det[b-a, c-a, d-a]: +1 when d lies on the side of plane abc that
// the right-hand normal (b-a)x(c-a) points to.
int orient_sign(const float* a, const float* b, const float* c, const float* d);

// +1 iff e lies strictly inside the circumsphere of abcd, where
// orient_sign(a,b,c,d) > 0; 0 on the sphere.
int insphere_sign(const float* a, const float* b, const float* c, const float* d, const float* e);

}  // namespace tsg
