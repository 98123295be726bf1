// Device-resident BDF / Newton-GMRES time integrator (see nsim_integrator.h).
#include "nsim_integrator.h"
