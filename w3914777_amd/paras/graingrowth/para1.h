// Grain-growth parameter set 1 (values of the reference's graingrowth/paras/para1.h).
const highprecision dtime=0.005,mobil=5.0,dx=0.5,dy=0.5,dxdy=dx*dy,grcoef=0.1;
const float R=50;
const float Ry=dimY/2,Rx=dimX/2;
