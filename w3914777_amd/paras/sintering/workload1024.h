// Sintering workload: 1024^2, two touching radius-150 particles with tanh
// interfaces (width 4 cells) plus +-1e-3 noise, A=16, B=1, kappa_c=5,
// kappa_eta=2, Dvol=0.04, Dvap=0.002, Dsurf=16, Dgb=1.6, L=10, dt=1e-4,
// 5000 steps, tile 32.
typedef double highprecision;
int timesteps=5000;
const int dimX=1024,dimY=1024;
const int unitx=32,unity=32;
const highprecision threshold=0.05;
#define AMSTENCIL
#define MonitorGradient
#define INIT_CONFIG
const highprecision coefm=5.0,coefk=2.0,coefl=10.0,dvol=0.04,dvap=0.002,dsur=16,dgrb=1.6,dx=0.5,dy=0.5,dxdy=dx*dy,dtime=1.0e-4;
const highprecision A=16.0,B=1.0,width=4.0,noise=1.0e-3;
float R1=150,R2=150;
float Rx1=dimX/2-R1,Rx2=dimX/2+R1,Ry1=dimY/2,Ry2=dimY/2;
