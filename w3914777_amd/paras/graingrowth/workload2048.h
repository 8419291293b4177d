// Grain-growth workload: 2048^2, Q=36 order parameters seeded U[-1e-3,1e-3],
// dx=0.5, dt=0.005, L=5, kappa=0.5, alpha=beta=gamma=1, periodic, tile 32,
// per-field gradient monitor, 10000 steps.
typedef double highprecision;
int timesteps=10000;
const int dimX=2048,dimY=2048;
const int unitx=32,unity=32;
const int Q=36;
const int monitor_field=-1;
const highprecision threshold=0.05;
#define AMSTENCIL
#define MonitorGradient
#define INIT_CONFIG
const highprecision dtime=0.005,mobil=5.0,dx=0.5,dy=0.5,dxdy=dx*dy,grcoef=0.5;
const highprecision alpha=1.0,beta=1.0,gamma=1.0,amp=1.0e-3;
