/*
 * main.c -- tsg_driver, the command-line driver of the reference
 * (src/main.cu:15-638).  The whole flow lives in the library as
 * tsg_driver_main (host/driver.c), so tests can run it in-process.
 */
#include "../../include/tilespgemm.h"

int main(int argc, char **argv) { return tsg_driver_main(argc, argv); }
