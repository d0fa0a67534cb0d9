/* Problem configuration for building the PLUTO reference hot path as an
 * oracle: ideal MHD, Cartesian, 2D, PLM (MC limiter on every variable),
 * RK2 unsplit, constrained transport with the default UCT_HLL EMF average.
 * This is the per-problem configuration a PLUTO user supplies; it is input
 * to the reference build, not reference source. */
#define  PHYSICS                        MHD
#define  DIMENSIONS                     2
#define  COMPONENTS                     2
#define  GEOMETRY                       CARTESIAN
#define  BODY_FORCE                     NO
#define  FORCED_TURB                    NO
#define  COOLING                        NO
#define  RECONSTRUCTION                 LINEAR
#define  TIME_STEPPING                  RK2
#define  DIMENSIONAL_SPLITTING          NO
#define  NTRACER                        0
#define  USER_DEF_PARAMETERS            0
#define  EOS                            IDEAL
#define  ENTROPY_SWITCH                 NO
#define  DIVB_CONTROL                   CONSTRAINED_TRANSPORT
#define  BACKGROUND_FIELD               NO
#define  AMBIPOLAR_DIFFUSION            NO
#define  RESISTIVITY                    NO
#define  HALL_MHD                       NO
#define  THERMAL_CONDUCTION             NO
#define  VISCOSITY                      NO
#define  ROTATING_FRAME                 NO
#define  LIMITER                        MC_LIM
#define  ASSIGN_VECTOR_POTENTIAL        NO
