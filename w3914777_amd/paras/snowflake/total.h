// Run settings of the snowflake drivers (the reference's snowflake/paras/total.h).
typedef double highprecision;
int timesteps=500;
const int dimX=512,dimY=512;
const int unitx=16,unity=16;
const highprecision threshold=0.99;
const highprecision ratio=1.0;
#define AMSTENCIL
#define End2end
#define Monitor2
#define pi 3.1415926
