/* The 2D oracle configuration with the ARITHMETIC EMF average
 * (problem configuration input to the reference build). */
#include "../cfg2d/definitions.h"
#define  CT_EMF_AVERAGE                 ARITHMETIC
