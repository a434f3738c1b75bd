!> iso_c_binding interface of libafid_io (include/afid_io.h): the solver's
!! HDF5 files in the reference's layout.  Strings go to C null-terminated
!! through cstr().
module afid_io
    use iso_c_binding
    implicit none

    interface
        type(c_ptr) function afid_h5_last_error() bind(C, name='afid_h5_last_error')
            import
        end function
        !> HdfCreateBlankFile (HdfRoutines.F90:11-22)
        integer(c_int) function afid_h5_create_file(path) bind(C, name='afid_h5_create_file')
            import; character(kind=c_char), intent(in) :: path(*)
        end function
        !> h5gcreate_f of HdfCreateMeansFile (CalcMeanProfiles.F90:289-334)
        integer(c_int) function afid_h5_create_group(path, name) bind(C, name='afid_h5_create_group')
            import; character(kind=c_char), intent(in) :: path(*), name(*)
        end function
        !> HdfWriteRealHalo3D (HdfRoutines.F90:428-520)
        integer(c_int) function afid_h5_write_field3d(path, data, nx, ny, nz) bind(C, name='afid_h5_write_field3d')
            import; character(kind=c_char), intent(in) :: path(*); real(c_double), intent(in) :: data(*)
            integer(c_int64_t), value :: nx, ny, nz
        end function
        !> HdfReadContinua (HdfReadContinua.F90:11-107)
        integer(c_int) function afid_h5_read_field3d(path, data, nx, ny, nz) bind(C, name='afid_h5_read_field3d')
            import; character(kind=c_char), intent(in) :: path(*); real(c_double), intent(out) :: data(*)
            integer(c_int64_t), value :: nx, ny, nz
        end function
        integer(c_int) function afid_h5_dataset_dims(path, name, dims, maxrank) bind(C, name='afid_h5_dataset_dims')
            import; character(kind=c_char), intent(in) :: path(*), name(*)
            integer(c_int64_t), intent(out) :: dims(*); integer(c_int), value :: maxrank
        end function
        !> HdfSerialWriteIntScalar / HdfSerialWriteRealScalar (HdfRoutines.F90:25-96)
        integer(c_int) function afid_h5_write_int_scalar(path, name, v) bind(C, name='afid_h5_write_int_scalar')
            import; character(kind=c_char), intent(in) :: path(*), name(*); integer(c_int32_t), value :: v
        end function
        integer(c_int) function afid_h5_write_real_scalar(path, name, v) bind(C, name='afid_h5_write_real_scalar')
            import; character(kind=c_char), intent(in) :: path(*), name(*); real(c_double), value :: v
        end function
        !> HdfSerialReadIntScalar / HdfSerialReadRealScalar (HdfRoutines.F90:257-309)
        integer(c_int) function afid_h5_read_int_scalar(path, name, v) bind(C, name='afid_h5_read_int_scalar')
            import; character(kind=c_char), intent(in) :: path(*), name(*); integer(c_int32_t), intent(out) :: v
        end function
        integer(c_int) function afid_h5_read_real_scalar(path, name, v) bind(C, name='afid_h5_read_real_scalar')
            import; character(kind=c_char), intent(in) :: path(*), name(*); real(c_double), intent(out) :: v
        end function
        !> HdfSerialWriteReal1D (HdfRoutines.F90:139-177) / HdfSerialReadReal1D (:311-336)
        integer(c_int) function afid_h5_write_real1d(path, name, data, n) bind(C, name='afid_h5_write_real1d')
            import; character(kind=c_char), intent(in) :: path(*), name(*); real(c_double), intent(in) :: data(*)
            integer(c_int64_t), value :: n
        end function
        integer(c_int) function afid_h5_read_real1d(path, name, data, n) bind(C, name='afid_h5_read_real1d')
            import; character(kind=c_char), intent(in) :: path(*), name(*); real(c_double), intent(out) :: data(*)
            integer(c_int64_t), value :: n
        end function
    end interface

contains

    !> trimmed Fortran string -> null-terminated C string
    function cstr(s) result(c)
        character(len=*), intent(in) :: s
        character(kind=c_char, len=len_trim(s) + 1) :: c
        c = trim(s) // c_null_char
    end function

    !> stop with the library's message if rc /= 0
    subroutine h5_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer :: i
        if (rc == 0) return
        call c_f_pointer(afid_h5_last_error(), msg, [512])
        text = ' '
        do i = 1, 512
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
        end do
        write(0, '(a,a,i0,a,a)') trim(what), ' failed (', rc, '): ', trim(text)
        error stop 1
    end subroutine
end module afid_io
