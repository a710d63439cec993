/* Test infrastructure only (CPU baseline): placeholder translation unit, the
 * calc2 port is added with the bench's cpu_baseline leg. */
int oracle_port_version(void) { return 1; }
