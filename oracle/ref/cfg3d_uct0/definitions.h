/* The 3D oracle configuration with the UCT0 EMF average
 * (problem configuration input to the reference build). */
#include "../cfg3d/definitions.h"
#define  CT_EMF_AVERAGE                 UCT0
