/* The 3D oracle configuration with the staggered field made from a vector
 * potential at start-up (problem configuration input to the reference build). */
#include "../cfg3d/definitions.h"
#undef   ASSIGN_VECTOR_POTENTIAL
#define  ASSIGN_VECTOR_POTENTIAL        YES
