// Snowflake workload: 4096^2, zero-flux boundaries, 5-point Laplacian,
// eps = 0.01(1+0.05cos(6(theta-pi/2))), tau=3e-4, alpha=0.9, gamma=10, K=1.6,
// Teq=1, radius-5 disc at the centre, interface noise a=0.01, tile 32,
// gradient monitor |grad phi| dx > 0.05, 20000 steps.
typedef double highprecision;
int timesteps=20000;
const int dimX=4096,dimY=4096;
const int unitx=32,unity=32;
const highprecision threshold=0.05;
#define AMSTENCIL
#define MonitorGradient
#define NEUMANN
#define LAP5
#define pi 3.1415926
const highprecision dx=0.03,dy=0.03,dxdy=dx*dy;
const highprecision tau = 0.0003;
const highprecision epsilonb = 0.01;
const highprecision kappa = 1.6;
const highprecision delta = 0.05;
const highprecision aniso = 6.0;
const highprecision alpha = 0.9;
const highprecision gama = 10.0;
const highprecision teq = 1.0;
const highprecision theta0 = pi/2;
const highprecision dtime=1.0e-4;
const highprecision noise=0.01;
const int seed=25;
const int cx=dimX/2,cy=dimY/2;
