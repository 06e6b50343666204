# Generates zerork_conf.h from the reference's own template
# (src/zerork/zerork_conf.h.in) with cmake's configure_file, exactly as the
# reference build does, selecting the libc exp (ZERORK_EXP_LIBC) so that the
# reference rates use the same exp() as a plain C restatement.
#
#   cmake -DZRK_SRC=<ref>/src/zerork -DZRK_OUT=<dir> -P ref_configure.cmake
set(ZERORK_EXP_LIBC 1)
configure_file(${ZRK_SRC}/zerork_conf.h.in ${ZRK_OUT}/zerork_conf.h)
