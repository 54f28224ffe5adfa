/* Force-driven channel, D3Q19, walls at y (no-slip), streaming method */
#define LB
#define LB_FLUID
#define LB_FLUID_FORCING
#define LB_FLUID_FORCING_POISEUILLE
#define LB_FLUID_BC
#define LB_FLUID_BC_Y  /* default fluid bc is no-slip */
#define GRID_POP_D3Q19
#define METHOD_STREAMING
#define METHOD_STEPPING_EULER
#define METHOD_FORCING_GUO
#define OUTPUT_AT_START
/* #define LB_TEMPERATURE */
// #define LB_FLUID_BC_X
