;;;; hgx.lisp
;;;;
;;;; sb-alien shim that routes Hypergeometrica's NTT multiply to libhgx, the
;;;; MI355X exact NTT multiply (C ABI: include/hgx.h).
;;;;
;;;; Load it after the HYPERGEOMETRICA system (as src/hgx.lisp at the end of
;;;; the system's file list, or from the REPL):
;;;;
;;;;   (setf hypergeometrica::*hgx-library* "/path/to/hypergeometrica_amd/libhgx.so")
;;;;   (load "lisp/hgx.lisp")
;;;;
;;;; It redefines the two functions the multiply dispatcher calls for
;;;; operands above *NTT-MULTIPLY-THRESHOLD* (src/multiply.lisp:12-49):
;;;;
;;;;   MPZ-*/NTT  (src/ntt-multiply.lisp:174-254)  ->  hgx_mul
;;;;   MPZ-SQUARE (src/ntt-multiply.lisp:101-172)  ->  hgx_sqr
;;;;
;;;; Their callers (MPD-*, binary splitting, MPZ-DIVREM, MPZ-EXPT) do not
;;;; change.  Beside them it offers what has no counterpart in the
;;;; reference:
;;;;
;;;;   MPZ-*-BATCH      one binary-splitting level in one call (hgx_mul_batch)
;;;;   MPZ-*-SUBMIT     asynchronous products through the request server
;;;;                    (hgx_submit / hgx_poll / hgx_wait)
;;;;   DMPZ             magnitudes held in device memory (hgx_device_alloc
;;;;                    and friends), for Newton iterations whose iterates
;;;;                    stay in HBM between products and sums
;;;;   HGX-BINARY-SPLIT BINARY-SPLIT of a series with polynomial terms by the
;;;;                    library's native engine (hgx_series_split): the same
;;;;                    PARTIAL, each tree height's products batched
;;;;
;;;; Limbs are the reference's storage: little-endian 64-bit digits in C
;;;; memory (src/vec.lisp:17-65, src/ram-vec.lisp:31-97); VEC-DIGIT-POINTER
;;;; gives their address as a system-area pointer, which is passed to the
;;;; library without copying.  Device pointers are plain addresses
;;;; (unsigned 64-bit integers) on the Lisp side.  Every foreign call runs
;;;; with floating-point traps masked: SBCL enables traps that the HIP
;;;; runtime may trip.

(in-package #:hypergeometrica)

;;; Loading

(defvar *hgx-library* "libhgx.so"
  "Path of libhgx.so (hypergeometrica_amd/libhgx.so in the MI355X build).")

(sb-alien:load-shared-object *hgx-library*)

(defmacro with-hgx-traps-masked (&body body)
  `(sb-int:with-float-traps-masked (:overflow :invalid :divide-by-zero :inexact)
     ,@body))

;;; Return codes (enum in include/hgx.h)

(defconstant +hgx-ok+ 0)
(defconstant +hgx-einval+ -1)
(defconstant +hgx-etoobig+ -2)
(defconstant +hgx-enomem+ -3)
(defconstant +hgx-edevice+ -4)
(defconstant +hgx-enotinit+ -5)

;;; Foreign types

;; One product for hgx_mul_batch / hgx_submit: OUT receives NA + NB limbs.
(sb-alien:define-alien-type hgx-mul-desc
    (sb-alien:struct hgx-mul-desc
      (a sb-sys:system-area-pointer)
      (na sb-alien:unsigned-long)
      (b sb-sys:system-area-pointer)
      (nb sb-alien:unsigned-long)
      (out sb-sys:system-area-pointer)))

;; A series for hgx_series_split: a(k), b(k), p(k), q(k) (index 0..3) as
;; polynomials in k, coefficients constant term first (COEF is the C array
;; coef[4][8], row by row), or VALUE0 at k = 0 where AT-ZERO is non-zero.
(sb-alien:define-alien-type hgx-series
    (sb-alien:struct hgx-series
      (degree (sb-alien:array sb-alien:int 4))
      (coef (sb-alien:array (sb-alien:signed 64) 32))
      (at-zero (sb-alien:array sb-alien:int 4))
      (value0 (sb-alien:array (sb-alien:signed 64) 4))))

;;; Foreign routines.  Pointers into host memory are system-area pointers;
;;; device pointers and streams are unsigned 64-bit addresses (0 = the
;;; library's stream).  Out-parameters come back as extra return values.

(sb-alien:define-alien-routine ("hgx_init" %hgx-init) sb-alien:int
  (device sb-alien:int) (budget sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_shutdown" %hgx-shutdown) sb-alien:void)
(sb-alien:define-alien-routine ("hgx_last_error" %hgx-last-error) sb-alien:c-string)
(sb-alien:define-alien-routine ("hgx_num_moduli" %hgx-num-moduli) sb-alien:int)
(sb-alien:define-alien-routine ("hgx_modulus" %hgx-modulus) sb-alien:unsigned-long
  (k sb-alien:int))
(sb-alien:define-alien-routine ("hgx_max_transform_bits" %hgx-max-transform-bits) sb-alien:int)

(sb-alien:define-alien-routine ("hgx_mul" %hgx-mul) sb-alien:int
  (a sb-sys:system-area-pointer) (na sb-alien:unsigned-long)
  (b sb-sys:system-area-pointer) (nb sb-alien:unsigned-long)
  (out sb-sys:system-area-pointer))
(sb-alien:define-alien-routine ("hgx_sqr" %hgx-sqr) sb-alien:int
  (a sb-sys:system-area-pointer) (na sb-alien:unsigned-long)
  (out sb-sys:system-area-pointer))
(sb-alien:define-alien-routine ("hgx_mul_batch" %hgx-mul-batch) sb-alien:int
  (descs (* hgx-mul-desc)) (count sb-alien:unsigned-long))

(sb-alien:define-alien-routine ("hgx_ntt_forward" %hgx-ntt-forward) sb-alien:int
  (data sb-sys:system-area-pointer) (n sb-alien:unsigned-long) (mod-num sb-alien:int))
(sb-alien:define-alien-routine ("hgx_ntt_reverse" %hgx-ntt-reverse) sb-alien:int
  (data sb-sys:system-area-pointer) (n sb-alien:unsigned-long) (mod-num sb-alien:int))

(sb-alien:define-alien-routine ("hgx_submit" %hgx-submit) sb-alien:int
  (desc (* hgx-mul-desc)) (ticket sb-alien:unsigned-long :out))
(sb-alien:define-alien-routine ("hgx_poll" %hgx-poll) sb-alien:int
  (ticket sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_wait" %hgx-wait) sb-alien:int
  (ticket sb-alien:unsigned-long))

(sb-alien:define-alien-routine ("hgx_device_alloc" %hgx-device-alloc) sb-alien:int
  (nlimbs sb-alien:unsigned-long) (d-out sb-alien:unsigned-long :out))
(sb-alien:define-alien-routine ("hgx_device_free" %hgx-device-free) sb-alien:int
  (d sb-alien:unsigned-long) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_upload" %hgx-upload) sb-alien:int
  (d-dst sb-alien:unsigned-long) (src sb-sys:system-area-pointer)
  (nlimbs sb-alien:unsigned-long) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_download" %hgx-download) sb-alien:int
  (dst sb-sys:system-area-pointer) (d-src sb-alien:unsigned-long)
  (nlimbs sb-alien:unsigned-long) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_device_copy" %hgx-device-copy) sb-alien:int
  (d-dst sb-alien:unsigned-long) (d-src sb-alien:unsigned-long)
  (nlimbs sb-alien:unsigned-long) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_device_fill" %hgx-device-fill) sb-alien:int
  (d sb-alien:unsigned-long) (nlimbs sb-alien:unsigned-long)
  (value sb-alien:unsigned-long) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_device_significant" %hgx-device-significant) sb-alien:int
  (d sb-alien:unsigned-long) (nlimbs sb-alien:unsigned-long)
  (out sb-alien:unsigned-long :out) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_device_compare" %hgx-device-compare) sb-alien:int
  (d-a sb-alien:unsigned-long) (na sb-alien:unsigned-long)
  (d-b sb-alien:unsigned-long) (nb sb-alien:unsigned-long)
  (out sb-alien:int :out) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_shl_device" %hgx-shl-device) sb-alien:int
  (d-a sb-alien:unsigned-long) (n sb-alien:unsigned-long) (bits sb-alien:int)
  (d-out sb-alien:unsigned-long) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_mul_device" %hgx-mul-device) sb-alien:int
  (d-a sb-alien:unsigned-long) (na sb-alien:unsigned-long)
  (d-b sb-alien:unsigned-long) (nb sb-alien:unsigned-long)
  (d-out sb-alien:unsigned-long) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_sqr_device" %hgx-sqr-device) sb-alien:int
  (d-a sb-alien:unsigned-long) (na sb-alien:unsigned-long)
  (d-out sb-alien:unsigned-long) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_add_device" %hgx-add-device) sb-alien:int
  (d-a sb-alien:unsigned-long) (na sb-alien:unsigned-long)
  (d-b sb-alien:unsigned-long) (nb sb-alien:unsigned-long)
  (d-out sb-alien:unsigned-long) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_sub_device" %hgx-sub-device) sb-alien:int
  (d-a sb-alien:unsigned-long) (na sb-alien:unsigned-long)
  (d-b sb-alien:unsigned-long) (nb sb-alien:unsigned-long)
  (d-out sb-alien:unsigned-long) (stream sb-alien:unsigned-long))
(sb-alien:define-alien-routine ("hgx_synchronize" %hgx-synchronize) sb-alien:int
  (stream sb-alien:unsigned-long))

(sb-alien:define-alien-routine ("hgx_autotune" %hgx-autotune) sb-alien:int
  (min-bits sb-alien:int) (max-bits sb-alien:int) (reps sb-alien:int))

;; the result of hgx_series_split is an opaque address (0 = none)
(sb-alien:define-alien-routine ("hgx_series_split" %hgx-series-split) sb-alien:int
  (series (* hgx-series)) (lower sb-alien:unsigned-long) (upper sb-alien:unsigned-long)
  (out sb-alien:unsigned-long :out))
(sb-alien:define-alien-routine ("hgx_partial_get" %hgx-partial-get) sb-alien:int
  (x sb-alien:unsigned-long) (which sb-alien:int) (sign sb-alien:int :out)
  (nlimbs sb-alien:unsigned-long :out) (limbs sb-sys:system-area-pointer))
(sb-alien:define-alien-routine ("hgx_partial_free" %hgx-partial-free) sb-alien:void
  (x sb-alien:unsigned-long))

;;; Errors

(define-condition hgx-error (error)
  ((code :initarg :code :reader hgx-error-code)
   (message :initarg :message :reader hgx-error-message))
  (:report (lambda (c s)
             (format s "hgx: ~A (code ~D)" (hgx-error-message c) (hgx-error-code c)))))

(defun hgx-check (rc)
  "Signal HGX-ERROR for a negative return code; return RC otherwise."
  (when (minusp rc)
    (error 'hgx-error :code rc :message (%hgx-last-error)))
  rc)

;;; Set-up

(defvar *hgx-ready* nil)
(defvar *hgx-device* 0
  "The GPU the library runs on.")
(defvar *hgx-hbm-budget* 0
  "Device work-memory cap in bytes (0: 90% of the free memory).")
(defvar *hgx-lock* (sb-thread:make-mutex :name "hgx set-up"))

(defun hgx-ensure ()
  "Initialise the library once: select the device and build the resident
prime and root tables (what the load-time construction of **SCHEME** does,
src/ntt-multiply.lisp:8, src/moduli.lisp:298-353)."
  (unless *hgx-ready*
    (sb-thread:with-mutex (*hgx-lock*)
      (unless *hgx-ready*
        (with-hgx-traps-masked
          (hgx-check (%hgx-init *hgx-device* *hgx-hbm-budget*)))
        (setf *hgx-ready* t))))
  t)

(defun hgx-shutdown ()
  "Release every device resource; the next product initialises again."
  (sb-thread:with-mutex (*hgx-lock*)
    (when *hgx-ready*
      (with-hgx-traps-masked (%hgx-shutdown))
      (setf *hgx-ready* nil))))

(defun hgx-autotune (&key (min-bits 13) (max-bits 27) (reps 2))
  "Time every pass structure and kernel shape per transform length and keep
the fastest (HGX_TUNE_CACHE=<file> in the environment keeps the choices
across processes)."
  (hgx-ensure)
  (with-hgx-traps-masked (hgx-check (%hgx-autotune min-bits max-bits reps))))

(defun hgx-moduli ()
  "The library's moduli, the ones of DEFAULT-MODULI (src/moduli.lisp:257-264)."
  (hgx-ensure)
  (loop :for k :below (%hgx-num-moduli) :collect (%hgx-modulus k)))

;;; The drop-in: MPZ-*/NTT and MPZ-SQUARE

(defun mpz-*/ntt (x y)
  "X * Y through libhgx.  Same contract as the reference's MPZ-*/NTT: a
fresh MPZ/RAM with sign (* (SIGN X) (SIGN Y)) whose storage holds the exact
product (NA + NB digits here; the dispatcher trims either way)."
  (hgx-ensure)
  (let* ((nx (mpz-size x))
         (ny (mpz-size y))
         (result (make-storage (+ nx ny))))
    (with-hgx-traps-masked
      (hgx-check (%hgx-mul (vec-digit-pointer (storage x)) nx
                           (vec-digit-pointer (storage y)) ny
                           (vec-digit-pointer result))))
    (make-instance 'mpz/ram :sign (* (sign x) (sign y)) :storage result)))

(defun mpz-square (x)
  "X * X through libhgx: a fresh MPZ/RAM of sign 1 (the reference's
MPZ-SQUARE)."
  (hgx-ensure)
  (let* ((n (mpz-size x))
         (result (make-storage (* 2 n))))
    (with-hgx-traps-masked
      (hgx-check (%hgx-sqr (vec-digit-pointer (storage x)) n
                           (vec-digit-pointer result))))
    (make-instance 'mpz/ram :sign 1 :storage result)))

;;; Transforms (NTT-FORWARD / NTT-REVERSE, src/number-theoretic-transform.lisp:40-143)

(defun hgx-ntt-forward (storage n mod-num)
  "Forward transform of N canonical residues in STORAGE (a ram-vec) modulo
the MOD-NUMth modulus, in place, output in bit-reversed order."
  (hgx-ensure)
  (with-hgx-traps-masked
    (hgx-check (%hgx-ntt-forward (vec-digit-pointer storage) n mod-num)))
  storage)

(defun hgx-ntt-reverse (storage n mod-num)
  "Inverse transform (bit-reversed input, natural-order output scaled by 1/N)."
  (hgx-ensure)
  (with-hgx-traps-masked
    (hgx-check (%hgx-ntt-reverse (vec-digit-pointer storage) n mod-num)))
  storage)

;;; One binary-splitting level in one call

(defun %fill-desc (desc x y out)
  (let ((same (eq x y)))
    (setf (sb-alien:slot desc 'a) (vec-digit-pointer (storage x))
          (sb-alien:slot desc 'na) (mpz-size x)
          (sb-alien:slot desc 'b) (vec-digit-pointer (storage (if same x y)))
          (sb-alien:slot desc 'nb) (mpz-size (if same x y))
          (sb-alien:slot desc 'out) (vec-digit-pointer out))))

(defun mpz-*-batch (pairs)
  "The products of PAIRS, a list of (X . Y) conses of MPZ/RAMs whose smaller
operand is above *NTT-MULTIPLY-THRESHOLD*, in one hgx_mul_batch call (the
products of one binary-splitting level, grouped by transform length on the
device).  A pair whose two operands are EQ is computed as a square.
Returns the list of fresh MPZ/RAM products, in order, each the value
MPZ-* gives."
  (hgx-ensure)
  (let* ((count (length pairs))
         (results (loop :for (x . y) :in pairs
                        :collect (make-storage (+ (mpz-size x) (mpz-size y)))))
         (descs (sb-alien:make-alien hgx-mul-desc (max count 1))))
    (unwind-protect
         (progn
           (loop :for (x . y) :in pairs
                 :for out :in results
                 :for i :from 0
                 :do (%fill-desc (sb-alien:deref descs i) x y out))
           (when (plusp count)
             (with-hgx-traps-masked
               (hgx-check (%hgx-mul-batch descs count)))))
      (sb-alien:free-alien descs))
    (loop :for (x . y) :in pairs
          :for out :in results
          :collect (optimize-mpz
                    (make-instance 'mpz/ram
                                   :sign (if (eq x y) 1 (* (sign x) (sign y)))
                                   :storage out)))))

;;; Asynchronous products through the request server

(defstruct (hgx-ticket (:constructor %make-hgx-ticket (id x y result sign desc)))
  id x y result sign desc (done nil))

;; Every pending ticket stays reachable here until it completes: the
;; operands' and the result's C memory is freed by the ram-vec finalizer
;; (src/ram-vec.lisp:95-96), and the server reads and writes it until then.
(defvar *hgx-pending* (make-hash-table :synchronized t))

(defun mpz-*-submit (x y)
  "Queue X * Y on the request server and return an HGX-TICKET at once.
Requests that arrive while others run are admitted as device memory frees
up, and their work overlaps on the device."
  (hgx-ensure)
  (let* ((result (make-storage (+ (mpz-size x) (mpz-size y))))
         (desc (sb-alien:make-alien hgx-mul-desc)))
    (%fill-desc (sb-alien:deref desc) x y result)
    (multiple-value-bind (rc id)
        (with-hgx-traps-masked (%hgx-submit desc))
      (when (minusp rc)
        (sb-alien:free-alien desc)
        (hgx-check rc))
      (let ((ticket (%make-hgx-ticket id x y result
                                      (if (eq x y) 1 (* (sign x) (sign y)))
                                      desc)))
        (setf (gethash id *hgx-pending*) ticket)
        ticket))))

(defun %hgx-ticket-finish (ticket rc)
  (setf (hgx-ticket-done ticket) t)
  (remhash (hgx-ticket-id ticket) *hgx-pending*)
  (sb-alien:free-alien (hgx-ticket-desc ticket))
  (hgx-check rc)
  (optimize-mpz (make-instance 'mpz/ram :sign (hgx-ticket-sign ticket)
                                        :storage (hgx-ticket-result ticket))))

(defun hgx-ticket-poll (ticket)
  "NIL while the product is pending; the product (an MPZ/RAM) once it is
done.  Signals HGX-ERROR if the request failed.  The library releases the
ticket when it reports completion, so poll a completed ticket only once."
  (assert (not (hgx-ticket-done ticket)) () "ticket already completed")
  (let ((rc (with-hgx-traps-masked (%hgx-poll (hgx-ticket-id ticket)))))
    (if (zerop rc)
        nil
        (%hgx-ticket-finish ticket (if (= rc 1) 0 rc)))))

(defun hgx-ticket-wait (ticket)
  "Block until the product is done and return it (an MPZ/RAM)."
  (assert (not (hgx-ticket-done ticket)) () "ticket already completed")
  (%hgx-ticket-finish ticket (with-hgx-traps-masked
                               (%hgx-wait (hgx-ticket-id ticket)))))

;;; Magnitudes in device memory (DMPZ)
;;;
;;; A DMPZ holds the digits of an MPZ in HBM: a device address, its digit
;;; count and a sign.  Products, squares, sums and differences of DMPZs run
;;; on the device, so a Newton iteration (MPD-RECIPROCAL,
;;; src/mpd-reciprocal.lisp:27-73; MPD-INV-SQRT, src/mpd-sqrt.lisp:40-87)
;;; keeps its iterates there and reads back only sizes and signs.  The
;;; device memory is released by DMPZ-FREE, or by a finalizer once the DMPZ
;;; is garbage.

(defstruct (dmpz (:constructor %make-dmpz (address size sign)))
  (address 0 :type (unsigned-byte 64))
  (size 0 :type (unsigned-byte 62))
  (sign 1 :type (member -1 1)))

(defun %dmpz-alloc (nlimbs sign)
  (multiple-value-bind (rc address)
      (with-hgx-traps-masked (%hgx-device-alloc (max nlimbs 1)))
    (hgx-check rc)
    (let ((d (%make-dmpz address nlimbs sign)))
      (sb-ext:finalize d (lambda ()
                           (when *hgx-ready*
                             (with-hgx-traps-masked (%hgx-device-free address 0))))
                       :dont-save t)
      d)))

(defun dmpz-free (d)
  "Release D's device memory now (after the work queued before it)."
  (sb-ext:cancel-finalization d)
  (when (plusp (dmpz-address d))
    (with-hgx-traps-masked (hgx-check (%hgx-device-free (dmpz-address d) 0)))
    (setf (dmpz-address d) 0 (dmpz-size d) 0)))

(defun %dmpz-trim (d)
  "Lower D's size to its significant digits (OPTIMIZE-MPZ,
src/mpz-ram.lisp:83-92), read from the device."
  (multiple-value-bind (rc n)
      (with-hgx-traps-masked
        (%hgx-device-significant (dmpz-address d) (dmpz-size d) 0))
    (hgx-check rc)
    (setf (dmpz-size d) n)
    d))

(defun mpz-dmpz (x)
  "Copy the MPZ/RAM X to device memory."
  (hgx-ensure)
  (let* ((n (mpz-size x))
         (d (%dmpz-alloc n (sign x))))
    (when (plusp n)
      (with-hgx-traps-masked
        (hgx-check (%hgx-upload (dmpz-address d) (vec-digit-pointer (storage x)) n 0))))
    d))

(defun dmpz-mpz (d)
  "Copy D back to a fresh MPZ/RAM."
  (let* ((n (dmpz-size d))
         (storage (make-storage (max n 1))))
    (when (plusp n)
      (with-hgx-traps-masked
        (hgx-check (%hgx-download (vec-digit-pointer storage) (dmpz-address d) n 0))))
    (optimize-mpz (make-instance 'mpz/ram :sign (dmpz-sign d) :storage storage))))

(defun dmpz-* (x y)
  "X * Y in device memory (a square when X and Y are EQ, the route
src/multiply.lisp:43-44 takes)."
  (let ((nx (dmpz-size x)) (ny (dmpz-size y)))
    (if (or (zerop nx) (zerop ny))
        (%dmpz-alloc 0 1)
        (let ((out (%dmpz-alloc (+ nx ny) (if (eq x y) 1 (* (dmpz-sign x) (dmpz-sign y))))))
          (with-hgx-traps-masked
            (hgx-check
             (if (eq x y)
                 (%hgx-sqr-device (dmpz-address x) nx (dmpz-address out) 0)
                 (%hgx-mul-device (dmpz-address x) nx (dmpz-address y) ny
                                  (dmpz-address out) 0))))
          (%dmpz-trim out)))))

(defun %dmpz-compare-magnitudes (x y)
  (multiple-value-bind (rc s)
      (with-hgx-traps-masked
        (%hgx-device-compare (dmpz-address x) (dmpz-size x)
                             (dmpz-address y) (dmpz-size y) 0))
    (hgx-check rc)
    s))

(defun %dmpz-add-magnitudes (x y sign)
  (let* ((n (max (dmpz-size x) (dmpz-size y)))
         (out (%dmpz-alloc (1+ n) sign)))
    (with-hgx-traps-masked
      (hgx-check (%hgx-add-device (dmpz-address x) (dmpz-size x)
                                  (dmpz-address y) (dmpz-size y)
                                  (dmpz-address out) 0)))
    (%dmpz-trim out)))

(defun %dmpz-sub-magnitudes (x y sign)
  "|X| - |Y| with |X| >= |Y|, of sign SIGN."
  (let ((out (%dmpz-alloc (dmpz-size x) sign)))
    (with-hgx-traps-masked
      (hgx-check (%hgx-sub-device (dmpz-address x) (dmpz-size x)
                                  (dmpz-address y) (dmpz-size y)
                                  (dmpz-address out) 0)))
    (%dmpz-trim out)))

(defun dmpz-+ (x y)
  "X + Y in device memory, with the sign rule of MPZ-+ on MPZ/RAM
(src/mpz-ram.lisp:185-291)."
  (cond
    ((zerop (dmpz-size y)) (%dmpz-copy x))
    ((zerop (dmpz-size x)) (%dmpz-copy y))
    ((= (dmpz-sign x) (dmpz-sign y))
     (%dmpz-add-magnitudes x y (dmpz-sign x)))
    (t
     (ecase (%dmpz-compare-magnitudes x y)
       (0 (%dmpz-alloc 0 1))
       (1 (%dmpz-sub-magnitudes x y (dmpz-sign x)))
       (-1 (%dmpz-sub-magnitudes y x (dmpz-sign y)))))))

(defun dmpz-negate (x)
  "-X, sharing nothing with X."
  (let ((d (%dmpz-copy x)))
    (unless (zerop (dmpz-size d))
      (setf (dmpz-sign d) (- (dmpz-sign x))))
    d))

(defun dmpz-- (x y)
  "X - Y in device memory."
  (let ((ny (dmpz-negate y)))
    (unwind-protect (dmpz-+ x ny)
      (dmpz-free ny))))

(defun %dmpz-copy (x)
  (let ((d (%dmpz-alloc (dmpz-size x) (dmpz-sign x))))
    (when (plusp (dmpz-size x))
      (with-hgx-traps-masked
        (hgx-check (%hgx-device-copy (dmpz-address d) (dmpz-address x) (dmpz-size x) 0))))
    d))

(defun dmpz-shift-left (x bits)
  "X * 2^BITS for 0 <= BITS < 64 (MPD-HALF! is the shift by 63 with the
exponent lowered by one digit, src/mpd.lisp:77-81)."
  (check-type bits (integer 0 63))
  (let ((out (%dmpz-alloc (1+ (dmpz-size x)) (dmpz-sign x))))
    (when (plusp (dmpz-size x))
      (with-hgx-traps-masked
        (hgx-check (%hgx-shl-device (dmpz-address x) (dmpz-size x) bits
                                    (dmpz-address out) 0))))
    (%dmpz-trim out)))

(defun dmpz-truncate (x keep)
  "The top KEEP digits of X as a fresh DMPZ, and the number of digits
dropped (MPD-TRUNCATE!'s mantissa step, src/mpd.lisp:56-70)."
  (let ((drop (max 0 (- (dmpz-size x) keep))))
    (if (zerop drop)
        (values (%dmpz-copy x) 0)
        (let ((d (%dmpz-alloc keep (dmpz-sign x))))
          (with-hgx-traps-masked
            (hgx-check (%hgx-device-copy (dmpz-address d)
                                         (+ (dmpz-address x) (* 8 drop))
                                         keep 0)))
          (values (%dmpz-trim d) drop)))))

;;; Binary splitting by the native engine
;;;
;;; BINARY-SPLIT (src/binary-splitting.lisp:123-133) takes a SERIES of four
;;; closures.  HGX-BINARY-SPLIT takes the same four term functions as
;;; polynomials, the form every series of the reference has, and returns
;;; the PARTIAL that BINARY-SPLIT returns for them: the library builds the
;;; same tree and folds its small subtrees.  A large split keeps its values
;;; in device memory -- the subtrees folded there, one thread each, each round
;;; of a tree height one batched device call and the sums on the device; a
;;; small one folds its subtrees on host threads, sends each height's
;;; products above *NTT-MULTIPLY-THRESHOLD* to the GPU in one batched call
;;; and multiplies the others on host threads.

(defun %fill-term (desc index term)
  "TERM is a list of integer coefficients, constant term first, optionally
preceded by :AT-ZERO V for a function whose value at k = 0 is V."
  (let ((at-zero nil) (v0 0) (coefs term))
    (when (eq (first term) :at-zero)
      (setf at-zero t v0 (second term) coefs (cddr term)))
    (assert (<= 1 (length coefs) 8) () "term polynomials have degree 0..7")
    (setf (sb-alien:deref (sb-alien:slot desc 'degree) index) (1- (length coefs))
          (sb-alien:deref (sb-alien:slot desc 'at-zero) index) (if at-zero 1 0)
          (sb-alien:deref (sb-alien:slot desc 'value0) index) v0)
    (loop :for j :below 8
          :for c := (or (nth j coefs) 0)
          :do (setf (sb-alien:deref (sb-alien:slot desc 'coef) (+ (* 8 index) j)) c))))

(defun %partial-mpz (handle which)
  (multiple-value-bind (rc sign n)
      (with-hgx-traps-masked (%hgx-partial-get handle which (sb-sys:int-sap 0)))
    (hgx-check rc)
    (let ((storage (make-storage (max n 1))))
      (when (plusp n)
        (multiple-value-bind (rc2 s2 n2)
            (with-hgx-traps-masked (%hgx-partial-get handle which (vec-digit-pointer storage)))
          (declare (ignore s2 n2))
          (hgx-check rc2)))
      (optimize-mpz (make-instance 'mpz/ram :sign (if (zerop sign) 1 sign) :storage storage)))))

(defun hgx-binary-split (a b p q lower upper)
  "The PARTIAL of BINARY-SPLIT over [LOWER, UPPER) for the series with term
functions A, B, P, Q given as polynomial lists (see %FILL-TERM)."
  (assert (> upper lower))
  (hgx-ensure)
  (let ((desc (sb-alien:make-alien hgx-series))
        (handle 0))
    (unwind-protect
         (progn
           (loop :for term :in (list a b p q)
                 :for i :from 0
                 :do (%fill-term (sb-alien:deref desc) i term))
           (multiple-value-bind (rc h)
               (with-hgx-traps-masked (%hgx-series-split desc lower upper))
             (hgx-check rc)
             (setf handle h))
           (make-partial :lower lower :upper upper
                         :p (%partial-mpz handle 0) :q (%partial-mpz handle 1)
                         :b (%partial-mpz handle 2) :r (%partial-mpz handle 3)))
      (unless (zerop handle) (with-hgx-traps-masked (%hgx-partial-free handle)))
      (sb-alien:free-alien desc))))

(defun hgx-chudnovsky-terms ()
  "MAKE-CHUDNOVSKY-SERIES's term functions (src/pi.lisp:24-43) as polynomials:
a, b, p, q for HGX-BINARY-SPLIT."
  (list (list (* 10 13591409) (* 10 545140134))
        (list (floor (* 8 640320) 12))
        (list :at-zero 1 5 -46 108 -72)
        (list :at-zero 1 0 0 0 (floor (expt 640320 3) 24))))
