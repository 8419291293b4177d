// Snowflake parameter set 3 (values of the reference's snowflake/paras/para3.h).
const highprecision dx=0.03,dy=0.03,dxdy=dx*dy;
const highprecision tau = 0.0003;
const highprecision epsilonb = 0.01;
const highprecision kappa = 1.8;
const highprecision delta = 0.02;
const highprecision aniso = 5.0;
const highprecision alpha = 0.9;
const highprecision gama = 10.0;
const highprecision teq = 1.0;
const highprecision theta0 = 0.0;
const highprecision dtime=1.0e-4;
const int seed=10;
