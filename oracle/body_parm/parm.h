/* Parameter file for the body() binding test: the reference's run-time
 * parameters are compile-time macros in its parm.h (TD_GPU_code/HighOrder/
 * parm.h:1-13), edited per run.  These are the same names at a grid small
 * enough for a test (1024 x 32 on rho in [-50, 50]); main.c fixes
 * dt = 0.002, a = 1e-8 and the m-loop -5..5 (main.c:40-55).  T = 260 runs
 * past body.cu's first written row (k dt >= Xmax + 200, body.cu:207), so
 * each waveform file gets about 100 rows. */
#define N 1024
#define M 32
#define Xmax 50.0
#define Xmin -50.0
#define Text 90.0
#define Nd 100
#define Fr 50
#define T 260.0

#define PTSX 128
#define PTSY 128
#define Pi 3.14159265358979
#define idx(b,a) (N*(b)+(a))
